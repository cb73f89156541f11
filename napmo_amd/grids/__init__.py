from napmo_amd.grids.angular import AngularGrid  # noqa: F401
from napmo_amd.grids.atomic import AtomicGrid  # noqa: F401
from napmo_amd.grids.becke import BeckeGrid  # noqa: F401
from napmo_amd.grids.cubic_spline import CubicSpline  # noqa: F401
from napmo_amd.grids.extrapolation import (CuspExtrapolation, PotentialExtrapolation,  # noqa: F401
                                           PowerExtrapolation, ZeroExtrapolation)
from napmo_amd.grids.lebedev import lebedev_get_order  # noqa: F401
from napmo_amd.grids.poisson_solver import poisson_solver  # noqa: F401
from napmo_amd.grids.radial import RadialGrid  # noqa: F401
from napmo_amd.grids.radial_transform import (ChebyshevRadialTransform,  # noqa: F401
                                              IndentityRadialTransform, PowerRadialTransform)
