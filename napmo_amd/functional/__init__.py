from napmo_amd.functional.interspecies_correlation import (  # noqa: F401
    isc_functional_selector, epc17_2, c_2species_grid)
