"""ctypes wrapper of the oracle (oracle/liboracle.so): the checker the tests compare against."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libffx_ref.so")

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so is missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.oracle_resolve_f32.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, fp, fp, fp, fp, ip, C.c_int, C.c_int]
        L.oracle_resolve_f64.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp, dp, fp, dp, dp, dp, ip, C.c_int, C.c_int]
        L.oracle_cas_f32.argtypes = [fp, fp, C.c_int, C.c_int, C.c_uint32, C.c_int]
        L.oracle_cas_f64.argtypes = [dp, dp, C.c_int, C.c_int, C.c_uint32, C.c_int]
        L.oracle_cas_setup.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)] + [C.c_float] * 5
        L.oracle_halton.restype = C.c_float
        L.oracle_halton.argtypes = [C.c_int, C.c_int]
        L.oracle_half_bits_trunc.restype = C.c_uint32
        L.oracle_half_bits_trunc.argtypes = [C.c_float]
        L.oracle_jitter.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, fp,
                                    C.c_uint32, C.c_uint32, fp, C.POINTER(C.c_uint32)]
        _L = L
    return _L


def ref_lib():
    """oracle/_ref/libffx_ref.so (the reference's own ffx_a.h/ffx_cas.h built for the CPU) or None."""
    if not os.path.exists(REF):
        return None
    L = C.CDLL(REF)
    L.ref_cas_setup.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)] + [C.c_float] * 5
    L.ref_half_bits.restype = C.c_uint32
    L.ref_half_bits.argtypes = [C.c_float]
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def resolve(U, color4, depth, vel4, hist4, hdepth, dtype=np.float32, cr16=True, threads=0, debug=True, inspect=True):
    """One frame of the oracle. Inputs are float32 arrays as the resolve reads them (vel4 already
    expanded to 4 channels); hist4 in `dtype`. The output size is the history's, the input size the
    colour's (they differ for TAAU). Returns dict(screen, history, debug, inspect)."""
    Hl, Wl = depth.shape
    H, W = hist4.shape[:2]
    threads = threads or os.cpu_count() or 1
    f = np.float32
    color4 = np.ascontiguousarray(color4, f)
    depth = np.ascontiguousarray(depth, f)
    vel4 = np.ascontiguousarray(vel4, f)
    hdepth = np.ascontiguousarray(hdepth, f)
    hist4 = np.ascontiguousarray(hist4, dtype)
    scr = np.empty((H, W, 4), dtype)
    his = np.empty((H, W, 4), dtype)
    dbg = np.empty((H, W, 4), dtype) if debug else None
    ins = np.empty((H, W, 16), np.int32) if inspect else None
    T = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_resolve_f32 if dtype == np.float32 else lib().oracle_resolve_f64
    rc = fn(C.byref(U), W, H, Wl, Hl, _p(color4, C.c_float), _p(depth, C.c_float), _p(vel4, C.c_float), _p(hist4, T),
            _p(hdepth, C.c_float), _p(scr, T), _p(his, T), _p(dbg, T) if debug else None,
            _p(ins, C.c_int32) if inspect else None, int(cr16), threads)
    assert rc == 0
    return dict(screen=scr, history=his, debug=dbg, inspect=ins)


def cas(src4, const1_x, dtype=np.float32, threads=0):
    H, W = src4.shape[:2]
    threads = threads or os.cpu_count() or 1
    src4 = np.ascontiguousarray(src4, dtype)
    dst = np.empty_like(src4)
    T = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().oracle_cas_f32 if dtype == np.float32 else lib().oracle_cas_f64
    assert fn(_p(src4, T), _p(dst, T), W, H, const1_x, threads) == 0
    return dst


def cas_setup(s, w, h):
    c0, c1 = (C.c_uint32 * 4)(), (C.c_uint32 * 4)()
    lib().oracle_cas_setup(c0, c1, s, w, h, w, h)
    return list(c0), list(c1)


def jitter(dist, frame, W, H, slow=0, fixed=-1, rot=0.0, scale=1.0, user=None):
    out = (C.c_float * 2)()
    n = C.c_uint32()
    u = None
    if user is not None:
        u = np.ascontiguousarray(np.asarray(user, np.float32).reshape(-1))
    rc = lib().oracle_jitter(dist, frame, slow, fixed, rot, scale, 0 if u is None else u.size // 2,
                             None if u is None else _p(u, C.c_float), W, H, out, C.byref(n))
    assert rc == 0
    return (out[0], out[1]), n.value


def velocity(R, depth, ids=None, objects=None, dtype=np.float32, threads=0):
    """Velocity image (H, W, 4) of the oracle's producer; objects: host float32 [n, 16] column-major."""
    H, W = depth.shape
    threads = threads or os.cpu_count() or 1
    L = lib()
    T = C.c_float if dtype == np.float32 else C.c_double
    fn = L.oracle_velocity_f32 if dtype == np.float32 else L.oracle_velocity_f64
    fn.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_int, C.c_int, C.POINTER(T), C.c_int]
    r = type(R).from_buffer_copy(R)
    obj = None if objects is None else np.ascontiguousarray(objects, np.float32)
    r.object_prev_from_current = None if obj is None else obj.ctypes.data
    depth = np.ascontiguousarray(depth, np.float32)
    ids_c = None if ids is None else np.ascontiguousarray(ids, np.uint32)
    out = np.empty((H, W, 4), dtype)
    rc = fn(C.byref(r), _p(depth, C.c_float), None if ids_c is None else _p(ids_c, C.c_uint32), W, H, _p(out, T),
            threads)
    assert rc == 0
    return out


def resolve_seg(U, color4, depth, vel4, hist4, hdepth, material, prev_material, uv_normal, prev_segmask,
                mask_init=None, dtype=np.float32, threads=0):
    """One frame of the oracle with the segmentation mask. The output size is the history's, the
    input size the depth's (material ids and G-buffer at the input size, masks at the output
    size). Returns the resolve dict plus `segmask`: mask_init (default zeros) with the texels the
    reference stores replaced."""
    Hl, Wl = depth.shape
    H, W = np.shape(hist4)[:2]
    threads = threads or os.cpu_count() or 1
    f, u = np.float32, np.uint32
    arrs = [np.ascontiguousarray(a, t) for a, t in ((color4, f), (depth, f), (vel4, f), (hdepth, f))]
    hist4 = np.ascontiguousarray(hist4, dtype)
    mats = [np.ascontiguousarray(a, u) for a in (material, prev_material, prev_segmask)]
    uvn = np.ascontiguousarray(uv_normal, f)
    mask = np.zeros((H, W), u) if mask_init is None else np.array(mask_init, u, copy=True)
    scr, his, dbg = (np.empty((H, W, 4), dtype) for _ in range(3))
    ins = np.empty((H, W, 16), np.int32)
    T = C.c_float if dtype == np.float32 else C.c_double
    L = lib()
    fn = L.oracle_resolve_seg_f32 if dtype == np.float32 else L.oracle_resolve_seg_f64
    fp, tp, up = C.POINTER(C.c_float), C.POINTER(T), C.POINTER(C.c_uint32)
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp, tp, fp, tp, tp, tp, C.POINTER(C.c_int32),
                   up, up, fp, up, up, C.c_int]
    rc = fn(C.byref(U), W, H, Wl, Hl, _p(arrs[0], C.c_float), _p(arrs[1], C.c_float), _p(arrs[2], C.c_float), _p(hist4, T),
            _p(arrs[3], C.c_float), _p(scr, T), _p(his, T), _p(dbg, T), _p(ins, C.c_int32), _p(mats[0], C.c_uint32),
            _p(mats[1], C.c_uint32), _p(uvn, C.c_float), _p(mats[2], C.c_uint32), _p(mask, C.c_uint32), threads)
    assert rc == 0
    return dict(screen=scr, history=his, debug=dbg, inspect=ins, segmask=mask)


def seg_sincos(a):
    a = np.ascontiguousarray(a, np.float32)
    s, c = np.empty_like(a), np.empty_like(a)
    L = lib()
    fp = C.POINTER(C.c_float)
    L.oracle_seg_sincos_f32.argtypes = [fp, fp, fp, C.c_int]
    L.oracle_seg_sincos_f32(_p(a, C.c_float), _p(s, C.c_float), _p(c, C.c_float), a.size)
    return s, c


def edge_aa(src4, mask, subpix=0.75, thr=0.116, thr_min=0.0833, dtype=np.float32, threads=0):
    """The edge-AA fallback over an RGBA image (H, W, 4) with its uint32 mask."""
    H, W = mask.shape
    threads = threads or os.cpu_count() or 1
    src = np.ascontiguousarray(src4, dtype)
    m = np.ascontiguousarray(mask, np.uint32)
    dst = np.empty_like(src)
    T = C.c_float if dtype == np.float32 else C.c_double
    L = lib()
    fn = L.oracle_edge_aa_f32 if dtype == np.float32 else L.oracle_edge_aa_f64
    fn.argtypes = [C.POINTER(T), C.POINTER(C.c_uint32), C.POINTER(T), C.c_int, C.c_int, T, T, T, C.c_int]
    assert fn(_p(src, T), _p(m, C.c_uint32), _p(dst, T), W, H, subpix, thr, thr_min, threads) == 0
    return dst
