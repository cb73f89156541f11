"""Launch-shape autotuner for grid Fock plans.

For a plan (one molecule's grid, basis and precision) it times every candidate launch
shape of the Fock contraction on the device (the per-tile FP64 / BF16x3 kernels' workgroup
count, the one-pass BF16x3 kernel's workgroup count, and the choice between the two
BF16x3 kernels) together with the K depth and prefetch depth of the density kernels' LDS
stages, then the synthesis' pairs per lane (1 or 2) together with the orbital-form density's
choice of kernel (C resident in LDS with Phi streamed in one of two launch shapes, or the
LDS-staged kernel), keeps the fastest, and stores it in a JSON cache keyed by the shape class
(precision, padded basis size, point count rounded to a power of two), so later plans of
the same class start from the cached choice without timing. The cache lives at
$NAPMO_TUNING_CACHE or ~/.cache/napmo_amd/tuning.json.
"""
import ctypes
import json
import math
import os
import time

from napmo_amd.system import cext

(TUNE_FOCK_WG, TUNE_PASS_WG, TUNE_USE_PASS, TUNE_DENSITY_KS, TUNE_DENSITY_PF, TUNE_SYNTH_NP,
 TUNE_DENSITY_STREAM, TUNE_INT8_CHUNK, TUNE_INT8_WAVE_BLOCKS) = range(9)
TUNE_INT8_PACK = 13
_KEYS = {'fock_wg': TUNE_FOCK_WG, 'pass_wg': TUNE_PASS_WG, 'use_pass': TUNE_USE_PASS,
         'density_ks': TUNE_DENSITY_KS, 'density_pf': TUNE_DENSITY_PF, 'synth_np': TUNE_SYNTH_NP,
         'density_stream': TUNE_DENSITY_STREAM, 'int8_chunk': TUNE_INT8_CHUNK,
         'int8_wave_blocks': TUNE_INT8_WAVE_BLOCKS, 'int8_pack': TUNE_INT8_PACK}
_INT8_KEYS = ('int8_chunk', 'int8_wave_blocks', 'int8_pack')   # only int8 plans have them
DEFAULTS = {'fock_wg': 0, 'pass_wg': 0, 'density_ks': 16, 'density_pf': 1,
            'synth_np': 2, 'density_stream': 1}   # 0: the library's default shape


def cache_path():
    return os.environ.get('NAPMO_TUNING_CACHE') or os.path.join(
        os.path.expanduser('~'), '.cache', 'napmo_amd', 'tuning.json')


def shape_key(fock):
    ld = -(-fock.nbasis // 64) * 64
    nb = 1 << max(0, int(round(math.log2(max(fock.npoint, 1)))))
    return '%s:ld%d:N%d' % (fock.precision, ld, nb)


def candidates(fock):
    """Launch shapes of the Fock contraction, each with both density stage depths."""
    if fock.precision == 'int8':
        # points per chunk (about N / 128 .. N / 512 chunks; every chunk is one workgroup per
        # row group and one FP64 partial of F; with screening, shorter chunks keep fewer
        # products but add partials to reduce) and the blocks per wave of the row-group packing
        N = fock.npoint
        chunks = sorted({min(4032, max(64, -(-N // k) + 63) // 64 * 64)
                         for k in (128, 192, 256, 320, 384, 448, 512)})
        out = [{'int8_chunk': c, 'int8_wave_blocks': b, 'int8_pack': 0} for c in chunks for b in (3, 4, 5)]
        # the kept blocks packed per chunk into tasks (the row-group packing then unused)
        out += [{'int8_chunk': c, 'int8_wave_blocks': 5, 'int8_pack': 1} for c in chunks]
    elif fock.precision == 'bf16x3':
        out = [{'use_pass': 1, 'pass_wg': w} for w in (128, 256, 512, 1024)]
        # the symmetric form only with short chunks (its FP32 chunk sums drift otherwise)
        out += [{'use_pass': 2, 'pass_wg': w} for w in (1024, 2048)]
        out += [{'use_pass': 0, 'fock_wg': w} for w in (1024, 2048, 4096)]
    else:
        out = [{'fock_wg': w} for w in (512, 1024, 2048, 4096, 8192)]
    ks = [16, 32] if (-(-fock.nbasis // 64) * 64) % 32 == 0 else [16]
    if fock.precision == 'int8':   # the orbital-form density's knobs are tuned in the second stage
        return out
    return [dict(c, density_ks=k, density_pf=f) for c in out for k in ks for f in (1, 2)]


def apply(fock, setting):
    for k, v in setting.items():
        if k in _KEYS and (k not in _INT8_KEYS or fock.precision == 'int8'):
            cext.check(cext.napmo_fock_plan_set_tuning(fock._this, _KEYS[k], int(v)))


def current(fock):
    out = {}
    for k, code in _KEYS.items():
        if k in _INT8_KEYS and fock.precision != 'int8':
            continue
        v = ctypes.c_long()
        cext.check(cext.napmo_fock_plan_get_tuning(fock._this, code, ctypes.byref(v)))
        out[k] = v.value
    return out


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, cache):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(fock, dP, dF, dE, reps=5, path=None, force=False, build=None, tag=''):
    """Choose and apply the fastest launch shape for ``fock`` (device buffers dP, dF, dE
    of the right sizes are used for the timing builds, or ``build()`` when given, e.g. an
    orbital-form build; ``tag`` then names that build in the cache key). Returns
    (setting, source) with source 'cache' or 'timed'."""
    path = path or cache_path()
    cache = _load(path)
    key = shape_key(fock) + (':' + tag if tag else '')
    run = build or (lambda: fock.build_device(dP, dF, dE))
    if key in cache and not force:
        setting = {k: v for k, v in cache[key].items() if k in _KEYS}
        try:
            apply(fock, dict(DEFAULTS, **setting))
            return setting, 'cache'
        except cext.NapmoError:
            pass   # not applicable to this plan: time again

    def timed(setting, n):
        apply(fock, dict(DEFAULTS, **setting))
        run()
        fock.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            run()
        fock.synchronize()
        return (time.perf_counter() - t0) / n

    scored = []
    for cand in candidates(fock):
        try:
            scored.append((cand, timed(cand, reps)))
        except cext.NapmoError:
            continue
    if not scored:
        raise cext.NapmoError('autotune: no launch shape applies to this plan')
    # the three fastest again with twice the builds (single timings differ by a few %)
    scored.sort(key=lambda r: r[1])
    best = min(((c, timed(c, 2 * reps)) for c, _ in scored[:3]), key=lambda r: r[1])
    per_np = [dict(best[0], synth_np=np_, density_stream=ds) for np_ in (1, 2) for ds in (0, 1, 2)]
    best = min(((c, timed(c, 2 * reps)) for c in per_np), key=lambda r: r[1])
    apply(fock, dict(DEFAULTS, **best[0]))
    cache[key] = dict(best[0], ms=best[1] * 1e3)
    _save(path, cache)
    return best[0], 'timed'
