"""ctypes front end of the CPU oracle (liborb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.  The
product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborb_oracle.so")
_lib = None

KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                           ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])


def build():
    """Compile liborb_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, I, F, V = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, None
        sig = {
            "oracle_extractor_create": (P, [I, F, I, I, I]),
            "oracle_extractor_destroy": (V, [P]),
            "oracle_extractor_tables": (I, [P, P, P, P, P, P, P]),
            "oracle_set_descriptor_variant": (I, [P, I]),
            "oracle_extract": (I, [P, P, I, I, I, P, I, P, P]),
            "oracle_level_size": (I, [P, I, P, P]),
            "oracle_level_image": (I, [P, I, I, P]),
            "oracle_level_candidates": (I, [P, I, P, I, P]),
            "oracle_resize_linear": (I, [P, I, I, I, P, I, I]),
            "oracle_distribute_octree": (I, [P, I, I, I, I, I, I, P, I, P]),
            "oracle_distribute_octree_leaves": (I, [P, I, I, I, I, I, I, P, I, P, P]),
            "oracle_plant": (I, [I, I]),
            "oracle_gaussian7": (I, [P, I, I, I, P]),
            "oracle_gaussian7_kernel": (I, [P]),
            "oracle_fast": (I, [P, I, I, I, I, P, I, P]),
            "oracle_fast_atan2": (F, [F, F]),
            "oracle_ic_angle": (F, [P, I, I, I]),
            "oracle_cos_sin": (V, [F, P, P]),
            "oracle_descriptor_distance": (I, [P, P]),
            "oracle_match_bf": (I, [P, I, P, I, I, F, P, P, I]),
            "oracle_knn2": (I, [P, I, P, ctypes.c_int64, F, P, P, P, I]),
            "oracle_stereo_match": (I, [P, P, P, P, I, P, P, I, F, F, P, P]),
            "oracle_stereo_frame": (I, [P, P, P, P, I, I, F, F, P, P, P, P, P, P, I, P, P]),
            "oracle_voc_create": (P, [I, I, I, I, I, P, P, P, P]),
            "oracle_voc_load_text": (P, [ctypes.c_char_p]),
            "oracle_voc_save_text": (I, [P, ctypes.c_char_p]),
            "oracle_voc_destroy": (V, [P]),
            "oracle_voc_info": (I, [P, P, P, P, P, P, P]),
            "oracle_voc_nodes": (I, [P, P, P, P, P]),
            "oracle_voc_transform": (I, [P, P, I, I, P, P, P]),
            "oracle_voc_bow": (I, [P, P, I, I, P, P, P, P, P, P]),
            "oracle_distinctive_descriptors": (I, [P, P, I, P]),
            "oracle_cvt_gray": (I, [P, I, I, I, I, I, P]),
            "oracle_init_rectify_map": (I, [P, P, I, P, P, I, I, P, P]),
            "oracle_remap_linear": (I, [P, I, I, I, P, P, I, I, P]),
            "oracle_undistort_points": (I, [P, I, P, P, I, P]),
            "oracle_rgbd_depth": (I, [P, P, I, P, I, I, I, F, P, P]),
            "oracle_search_for_initialization": (I, [P, P, I, P, P, I, P, P, I, F, I, P]),
            "oracle_search_by_projection": (I, [P, P, P, P, I, P, P, P, P, I, F, F, P, P]),
            "oracle_search_by_projection_last": (I, [P, P, P, P, I, P, P, P, I, P, P, I, F, I, P, P]),
            "oracle_planted": (I, [I]),
            "oracle_set_opencv_simd": (I, [I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class OracleExtractor:
    """ORBextractor(nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST) on the CPU."""

    def __init__(self, nfeatures=1000, scaleFactor=1.2, nlevels=8, iniThFAST=20, minThFAST=7):
        self.nfeatures, self.nlevels = nfeatures, nlevels
        self.h = lib().oracle_extractor_create(nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST)
        if not self.h:
            raise ValueError("invalid extractor parameters")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_extractor_destroy(self.h)
            self.h = None

    def tables(self):
        n = self.nlevels
        s, inv, s2, inv2 = (np.zeros(n, np.float32) for _ in range(4))
        quota = np.zeros(n, np.int32)
        umax = np.zeros(16, np.int32)
        lib().oracle_extractor_tables(self.h, _p(s), _p(inv), _p(s2), _p(inv2), _p(quota), _p(umax))
        return dict(scale=s, inv_scale=inv, sigma2=s2, inv_sigma2=inv2, quota=quota, umax=umax)

    def set_variant(self, v):
        if lib().oracle_set_descriptor_variant(self.h, v) != 0:
            raise ValueError(v)

    def capacity(self):
        return self.nfeatures + 3 * self.nlevels + 64

    def __call__(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        cap = self.capacity()
        kps = np.zeros(cap, KEYPOINT_DTYPE)
        desc = np.zeros((cap, 32), np.uint8)
        n = ctypes.c_int(0)
        rc = lib().oracle_extract(self.h, _p(img), img.shape[1], img.shape[0], img.shape[1], _p(kps), cap,
                                  _p(desc), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(f"oracle_extract failed: {rc}")
        return kps[:n.value].copy(), desc[:n.value].copy()

    def level_size(self, level):
        w, h = ctypes.c_int(), ctypes.c_int()
        lib().oracle_level_size(self.h, level, ctypes.byref(w), ctypes.byref(h))
        return w.value, h.value

    def level_image(self, level, blurred=False):
        w, h = self.level_size(level)
        out = np.zeros((h, w), np.uint8)
        lib().oracle_level_image(self.h, level, 1 if blurred else 0, _p(out))
        return out

    def candidates(self, level):
        cap = 1 << 20
        buf = np.zeros((cap, 3), np.int32)
        n = ctypes.c_int()
        rc = lib().oracle_level_candidates(self.h, level, _p(buf), cap, ctypes.byref(n))
        if rc != 0:
            raise RuntimeError(rc)
        return buf[:n.value].copy()


def distribute_octree(xyr, minX, maxX, minY, maxY, N):
    xyr = np.ascontiguousarray(xyr, np.int32).reshape(-1, 3)
    cap = len(xyr) + 64
    out = np.zeros((cap, 3), np.int32)
    n = ctypes.c_int()
    rc = lib().oracle_distribute_octree(_p(xyr), len(xyr), minX, maxX, minY, maxY, N, _p(out), cap, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(rc)
    return out[:n.value].copy()


def distribute_octree_leaves(xyr, minX, maxX, minY, maxY, N):
    """Kept nodes of DistributeOctTree: (n_leaves x 4) rects (ulx, uly, urx,
    bry) in list order, and the node index holding each input candidate."""
    xyr = np.ascontiguousarray(xyr, np.int32).reshape(-1, 3)
    cap = len(xyr) + 64
    rects = np.zeros((cap, 4), np.int32)
    leaf_of = np.full(len(xyr), -1, np.int32)
    n = ctypes.c_int()
    rc = lib().oracle_distribute_octree_leaves(_p(xyr), len(xyr), minX, maxX, minY, maxY, N, _p(rects), cap,
                                               _p(leaf_of), ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(rc)
    return rects[:n.value].copy(), leaf_of


PLANT_ATAN, PLANT_RETAIN, PLANT_STEREO, PLANT_PROJECTION = 1, 2, 3, 4

PROJ_POINT_DTYPE = np.dtype([("proj_x", "<f4"), ("proj_y", "<f4"), ("proj_xr", "<f4"), ("view_cos", "<f4"),
                             ("level", "<i4"), ("flags", "<i4")])
LAST_POINT_DTYPE = np.dtype([("xc", "<f4"), ("yc", "<f4"), ("zc", "<f4"), ("angle", "<f4"), ("octave", "<i4"),
                             ("flags", "<i4")])


def _frame_args(kp, desc, uright, occupied, bounds, scale):
    kp = np.ascontiguousarray(kp, KEYPOINT_DTYPE)
    desc = np.ascontiguousarray(desc, np.uint8).reshape(-1, 32)
    uright = np.ascontiguousarray(uright, np.float32)
    occupied = np.ascontiguousarray(occupied, np.uint8)
    bounds = np.ascontiguousarray(bounds, np.float32)
    scale = np.ascontiguousarray(scale, np.float32)
    assert len(desc) == len(kp) == len(uright) == len(occupied)
    return kp, desc, uright, occupied, bounds, scale


def search_by_projection(kp, desc, uright, occupied, bounds, scale, pts, pdesc, th=1.0, nnratio=0.8):
    """ORBmatcher::SearchByProjection(F, vpMapPoints, th) -> (nmatches,
    pt_match, key_point); see orb_oracle.h."""
    kp, desc, uright, occupied, bounds, scale = _frame_args(kp, desc, uright, occupied, bounds, scale)
    pts = np.ascontiguousarray(pts, PROJ_POINT_DTYPE)
    pdesc = np.ascontiguousarray(pdesc, np.uint8).reshape(-1, 32)
    pm = np.zeros(len(pts), np.int32)
    kpm = np.zeros(len(kp), np.int32)
    n = lib().oracle_search_by_projection(_p(kp), _p(desc), _p(uright), _p(occupied), len(kp), _p(bounds),
                                          _p(scale), _p(pts), _p(pdesc), len(pts), th, nnratio, _p(pm), _p(kpm))
    return n, pm, kpm


def search_by_projection_last(kp, desc, uright, occupied, bounds, scale, cam, mono, pts, pdesc, th,
                              check_orientation=True):
    """ORBmatcher::SearchByProjection(Current, Last, th, bMono) -> (nmatches,
    pt_match, key_point); cam = fx, fy, cx, cy, mbf, mb, tlc_z."""
    kp, desc, uright, occupied, bounds, scale = _frame_args(kp, desc, uright, occupied, bounds, scale)
    cam = np.ascontiguousarray(cam, np.float32)
    pts = np.ascontiguousarray(pts, LAST_POINT_DTYPE)
    pdesc = np.ascontiguousarray(pdesc, np.uint8).reshape(-1, 32)
    pm = np.zeros(len(pts), np.int32)
    kpm = np.zeros(len(kp), np.int32)
    n = lib().oracle_search_by_projection_last(_p(kp), _p(desc), _p(uright), _p(occupied), len(kp), _p(bounds),
                                               _p(scale), _p(cam), 1 if mono else 0, _p(pts), _p(pdesc), len(pts),
                                               th, 1 if check_orientation else 0, _p(pm), _p(kpm))
    return n, pm, kpm


class planted:
    """Context manager: run the oracle with a deliberate mistake planted in
    one function (see oracle_plant), restoring the faithful statement after."""

    def __init__(self, family, variant):
        self.family, self.variant = family, variant

    def __enter__(self):
        if lib().oracle_plant(self.family, self.variant) != 0:
            raise ValueError((self.family, self.variant))
        return self

    def __exit__(self, *exc):
        lib().oracle_plant(self.family, 0)
        return False


OPENCV_SSE2 = 3          # the default: both kernels as x86 OpenCV builds run them


class opencv_simd:
    """Context manager: which form of OpenCV's 8U kernels the oracle follows
    (flags bit 0: SSE2 resize vertical pass, bit 1: SSE2 Gaussian column
    pass; a clear bit selects the generic scalar pass).  Restores the
    default (both SSE2) on exit."""

    def __init__(self, flags):
        self.flags = flags

    def __enter__(self):
        if lib().oracle_set_opencv_simd(self.flags) != 0:
            raise ValueError(self.flags)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_opencv_simd(OPENCV_SSE2)
        return False


def resize_linear(src, dw, dh):
    src = np.ascontiguousarray(src, np.uint8)
    out = np.zeros((dh, dw), np.uint8)
    lib().oracle_resize_linear(_p(src), src.shape[1], src.shape[0], src.shape[1], _p(out), dw, dh)
    return out


def gaussian7(src):
    src = np.ascontiguousarray(src, np.uint8)
    out = np.zeros_like(src)
    if lib().oracle_gaussian7(_p(src), src.shape[1], src.shape[0], src.shape[1], _p(out)) != 0:
        raise ValueError("image too small")
    return out


def gaussian7_kernel():
    k = np.zeros(7, np.int32)
    lib().oracle_gaussian7_kernel(_p(k))
    return k


def fast(img, threshold):
    img = np.ascontiguousarray(img, np.uint8)
    cap = img.size
    buf = np.zeros((cap, 3), np.int32)
    n = ctypes.c_int()
    lib().oracle_fast(_p(img), img.shape[1], img.shape[0], img.shape[1], threshold, _p(buf), cap, ctypes.byref(n))
    return buf[:n.value].copy()


def fast_atan2(y, x):
    return lib().oracle_fast_atan2(y, x)


def ic_angle(img, cx, cy):
    img = np.ascontiguousarray(img, np.uint8)
    return lib().oracle_ic_angle(_p(img), img.shape[1], cx, cy)


def cos_sin(rad):
    c, s = ctypes.c_float(), ctypes.c_float()
    lib().oracle_cos_sin(rad, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


def descriptor_distance(a, b):
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    return lib().oracle_descriptor_distance(_p(a), _p(b))


def match_bf(q, t, th_low=50, ratio=0.6, nthreads=1):
    q = np.ascontiguousarray(q, np.uint8)
    t = np.ascontiguousarray(t, np.uint8)
    idx = np.zeros(len(q), np.int32)
    dist = np.zeros(len(q), np.int32)
    lib().oracle_match_bf(_p(q), len(q), _p(t), len(t), th_low, ratio, _p(idx), _p(dist), nthreads)
    return idx, dist


def knn2(q, m, ratio=0.7, nthreads=1):
    q = np.ascontiguousarray(q, np.uint8)
    m = np.ascontiguousarray(m, np.uint8)
    idx = np.zeros((len(q), 2), np.int32)
    dist = np.zeros((len(q), 2), np.int32)
    acc = np.zeros(len(q), np.uint8)
    lib().oracle_knn2(_p(q), len(q), _p(m), len(m), ratio, _p(idx), _p(dist), _p(acc), nthreads)
    return idx, dist, acc.astype(bool)


def stereo_match(left, right, kl, dl, kr, dr, bf, b):
    kl = np.ascontiguousarray(kl)
    kr = np.ascontiguousarray(kr)
    dl = np.ascontiguousarray(dl, np.uint8)
    dr = np.ascontiguousarray(dr, np.uint8)
    uR = np.zeros(len(kl), np.float32)
    dep = np.zeros(len(kl), np.float32)
    lib().oracle_stereo_match(left.h, right.h, _p(kl), _p(dl), len(kl), _p(kr), _p(dr), len(kr), bf, b,
                              _p(uR), _p(dep))
    return uR, dep


def stereo_frame(left, right, imL, imR, bf, b):
    """Frame-constructor order: L/R extraction on two threads, then stereo."""
    imL = np.ascontiguousarray(imL, np.uint8)
    imR = np.ascontiguousarray(imR, np.uint8)
    cap = left.capacity()
    kl = np.zeros(cap, KEYPOINT_DTYPE)
    kr = np.zeros(cap, KEYPOINT_DTYPE)
    dl = np.zeros((cap, 32), np.uint8)
    dr = np.zeros((cap, 32), np.uint8)
    nl, nr = ctypes.c_int(), ctypes.c_int()
    uR = np.zeros(cap, np.float32)
    dep = np.zeros(cap, np.float32)
    rc = lib().oracle_stereo_frame(left.h, right.h, _p(imL), _p(imR), imL.shape[1], imL.shape[0], bf, b,
                                   _p(kl), _p(dl), ctypes.byref(nl), _p(kr), _p(dr), ctypes.byref(nr), cap,
                                   _p(uR), _p(dep))
    if rc != 0:
        raise RuntimeError(rc)
    n = nl.value
    return kl[:n], dl[:n], kr[:nr.value], dr[:nr.value], uR[:n], dep[:n]


class OracleVocabulary:
    """DBoW2 TemplatedVocabulary restated (bow_oracle.cpp).  Node arrays are
    in file order with node 0 the root: parent[i] < i, is_word, desc (n x 32
    u8), weight (f64)."""

    def __init__(self, k=None, L=None, scoring=0, weighting=0, parent=None, is_word=None, desc=None,
                 weight=None, path=None):
        if path is not None:
            self.h = lib().oracle_voc_load_text(os.fsencode(path))
        else:
            parent = np.ascontiguousarray(parent, np.int32)
            n = len(parent)
            is_word = np.ascontiguousarray(is_word, np.uint8)
            desc = np.ascontiguousarray(desc, np.uint8).reshape(n, 32)
            weight = np.ascontiguousarray(weight, np.float64)
            self.h = lib().oracle_voc_create(k, L, scoring, weighting, n, _p(parent), _p(is_word), _p(desc),
                                             _p(weight))
        if not self.h:
            raise ValueError("vocabulary rejected")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_voc_destroy(self.h)
            self.h = None

    def info(self):
        v = [ctypes.c_int() for _ in range(6)]
        lib().oracle_voc_info(self.h, *[ctypes.byref(x) for x in v])
        return dict(zip(("k", "L", "scoring", "weighting", "n_nodes", "n_words"), (x.value for x in v)))

    def nodes(self):
        n = self.info()["n_nodes"]
        parent = np.zeros(n, np.int32)
        word = np.zeros(n, np.int32)
        desc = np.zeros((n, 32), np.uint8)
        weight = np.zeros(n, np.float64)
        lib().oracle_voc_nodes(self.h, _p(parent), _p(word), _p(desc), _p(weight))
        return parent, word, desc, weight

    def save_text(self, path):
        if lib().oracle_voc_save_text(self.h, os.fsencode(path)) != 0:
            raise OSError(path)

    def transform(self, desc, levelsup=0):
        desc = np.ascontiguousarray(desc, np.uint8).reshape(-1, 32)
        n = len(desc)
        wid = np.zeros(n, np.int32)
        w = np.zeros(n, np.float64)
        nid = np.zeros(n, np.int32)
        lib().oracle_voc_transform(self.h, _p(desc), n, levelsup, _p(wid), _p(w), _p(nid))
        return wid, w, nid

    def bow(self, desc, levelsup=0):
        desc = np.ascontiguousarray(desc, np.uint8).reshape(-1, 32)
        n = len(desc)
        bw = np.zeros(max(n, 1), np.int32)
        bv = np.zeros(max(n, 1), np.float64)
        fn = np.zeros(max(n, 1), np.int32)
        ff = np.zeros(max(n, 1), np.int32)
        nb, nf = ctypes.c_int(), ctypes.c_int()
        lib().oracle_voc_bow(self.h, _p(desc), n, levelsup, _p(bw), _p(bv), ctypes.byref(nb), _p(fn), _p(ff),
                             ctypes.byref(nf))
        return bw[:nb.value].copy(), bv[:nb.value].copy(), fn[:nf.value].copy(), ff[:nf.value].copy()


def distinctive_descriptors(points):
    """MapPoint::ComputeDistinctiveDescriptors for a list of N_p x 32 arrays:
    the chosen row of each (-1 for an empty one)."""
    mats = [np.ascontiguousarray(d, np.uint8).reshape(-1, 32) for d in points]
    offs = np.zeros(len(mats) + 1, np.int32)
    offs[1:] = np.cumsum([len(d) for d in mats])
    desc = np.concatenate(mats) if offs[-1] else np.zeros((1, 32), np.uint8)
    best = np.zeros(max(len(mats), 1), np.int32)
    lib().oracle_distinctive_descriptors(_p(desc), _p(offs), len(mats), _p(best))
    return best[:len(mats)]


def cvt_gray(img, rgb=True):
    """cvtColor(CV_RGB2GRAY / CV_BGR2GRAY / ...A2GRAY) of an h x w x 3|4 u8 image."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w, scn = img.shape
    out = np.zeros((h, w), np.uint8)
    lib().oracle_cvt_gray(_p(img), w, h, w * scn, scn, 2 if rgb else 0, _p(out))
    return out


def init_rectify_map(K, D, R, P, w, h):
    """initUndistortRectifyMap -> (mapx, mapy) float32 h x w."""
    K = np.ascontiguousarray(K, np.float64).reshape(9)
    D = np.ascontiguousarray(D if D is not None else [], np.float64).reshape(-1)
    R = None if R is None else np.ascontiguousarray(R, np.float64).reshape(9)
    P = np.ascontiguousarray(P, np.float64).reshape(3, -1)[:, :3].copy().reshape(9)
    mx = np.zeros((h, w), np.float32)
    my = np.zeros((h, w), np.float32)
    rc = lib().oracle_init_rectify_map(_p(K), _p(D) if len(D) else None, len(D), _p(R), _p(P), w, h, _p(mx), _p(my))
    if rc != 0:
        raise ValueError("bad rectification parameters")
    return mx, my


def remap_linear(src, mapx, mapy):
    """remap(src, mapx, mapy, INTER_LINEAR), border constant 0."""
    src = np.ascontiguousarray(src, np.uint8)
    mapx = np.ascontiguousarray(mapx, np.float32)
    mapy = np.ascontiguousarray(mapy, np.float32)
    h, w = mapx.shape
    out = np.zeros((h, w), np.uint8)
    lib().oracle_remap_linear(_p(src), src.shape[1], src.shape[0], src.shape[1], _p(mapx), _p(mapy), w, h, _p(out))
    return out


def search_for_initialization(kp1, d1, kp2, d2, bounds, prev, window=100, nnratio=0.9, check_orientation=True):
    """ORBmatcher::SearchForInitialization: (nmatches, matches12, updated prev)."""
    kp1 = np.ascontiguousarray(kp1, KEYPOINT_DTYPE)
    kp2 = np.ascontiguousarray(kp2, KEYPOINT_DTYPE)
    d1 = np.ascontiguousarray(d1, np.uint8).reshape(-1, 32)
    d2 = np.ascontiguousarray(d2, np.uint8).reshape(-1, 32)
    b = np.ascontiguousarray(bounds, np.float32)
    pv = np.array(prev, np.float32).reshape(-1, 2).copy()
    m = np.zeros(max(len(kp1), 1), np.int32)
    n = lib().oracle_search_for_initialization(_p(kp1), _p(d1), len(kp1), _p(kp2), _p(d2), len(kp2), _p(b), _p(pv),
                                               int(window), float(nnratio), int(bool(check_orientation)), _p(m))
    return n, m[:len(kp1)], pv


def undistort_points(xy, K, D):
    """Frame::UndistortKeyPoints' cv::undistortPoints (R = I, P = K) on an
    (n, 2) float32 array; K = (fx, fy, cx, cy), D = 4 or 5 coefficients."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    K = np.ascontiguousarray(K, np.float32)
    D = np.ascontiguousarray(D, np.float32)
    out = np.empty_like(xy)
    if lib().oracle_undistort_points(_p(xy), len(xy), _p(K), _p(D), len(D), _p(out)) != 0:
        raise ValueError("bad coefficients")
    return out


def rgbd_depth(xy, xun, depth, bf):
    """Frame::ComputeStereoFromRGBD: (uRight, depth) per keypoint."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    xun = np.ascontiguousarray(xun, np.float32)
    depth = np.ascontiguousarray(depth, np.float32)
    uR = np.empty(len(xy), np.float32)
    dep = np.empty(len(xy), np.float32)
    lib().oracle_rgbd_depth(_p(xy), _p(xun), len(xy), _p(depth), depth.shape[1], depth.shape[0], depth.shape[1],
                            float(bf), _p(uR), _p(dep))
    return uR, dep
