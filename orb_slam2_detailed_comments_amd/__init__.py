"""MI355X-native ORB front end for ORB-SLAM2's data-parallel hot path.

Python mirror of the reference interface over the C ABI in include/orb_gpu.h
(liborb_gpu.so, built in-tree by build.py):

* ``ORBextractor(nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST)``
  with ``__call__(image, mask=None) -> (keypoints, descriptors)``, the
  ``Get*`` accessors and ``mvImagePyramid``  (include/ORBextractor.h:93-162)
* ``ORBmatcher.DescriptorDistance(a, b)``     (include/ORBmatcher.h:65)
* ``match_bf``                                (best/second rule of
  ORBmatcher::SearchForInitialization, src/ORBmatcher.cc:614-647)
* ``compute_stereo_matches`` / ``stereo_frame`` (Frame::ComputeStereoMatches,
  src/Frame.cc:831-1082, and the stereo Frame constructor)
* ``distinctive_descriptors`` (MapPoint::ComputeDistinctiveDescriptors,
  src/MapPoint.cc:365-450, for many map points at once)
* ``Vocabulary`` with ``transform`` / ``bow`` (DBoW2 TemplatedVocabulary,
  Thirdparty/DBoW2/DBoW2/TemplatedVocabulary.h, as Frame::ComputeBoW uses it)

Keypoints are numpy structured arrays with cv::KeyPoint's fields.  There is no
CPU fallback: if the native library is missing every call raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["KEYPOINT_DTYPE", "Context", "ORBextractor", "ORBmatcher", "Batch", "Server", "Vocabulary", "match_bf",
           "undistort_keypoints", "rgbd_depth", "rgbd_frame",
           "knn2", "compute_stereo_matches", "stereo_frame", "distinctive_descriptors", "cvt_color_gray", "Rectifier", "search_for_initialization", "lib", "library_path",
           "OrbError"]

_PKG = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_PKG, "liborb_gpu.so")

KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                           ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])

ORB_OK, ORB_E_ARG, ORB_E_CAPACITY, ORB_E_DEVICE, ORB_E_OOM, ORB_E_OVERFLOW, ORB_E_ADMISSION = 0, -1, -2, -3, -4, -5, -6
_ERRORS = {-1: "bad argument", -2: "output capacity too small", -3: "HIP device error",
           -4: "device allocation failed", -5: "internal bound exceeded", -6: "refused by memory budget"}

# Every symbol the C ABI exports (include/orb_gpu.h), with its ctypes signature.
_P, _I, _F, _S, _V = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t, None
SIGNATURES = {
    "orb_context_create": (_I, [_I, _P]),
    "orb_batch_upload_images": (_I, [_P, _I, _P, _P, _P]),
    "orb_extractor_set_tile_heights": (_I, [_P, _I, _P]),
    "orb_context_set_round_limit": (_I, [_P, _I]),
    "orb_context_last_rounds": (_I, [_P, _P]),
    "orb_batch_set_tile_heights": (_I, [_P, _I, _P]),
    "orb_batch_autotune_tiles": (_I, [_P, _I, _P, _P, _P]),
    "orb_batch_read_all": (_I, [_P, _P, _P, _I, _P, _P, _P]),
    "orb_batch_pinned_images": (_I, [_P, _P, _P]),
    "orb_batch_upload_pinned": (_I, [_P]),
    "orb_batch_read_async": (_I, [_P]),
    "orb_batch_pinned_results": (_I, [_P, _I, _P, _P, _P]),
    "orb_batch_pinned_stereo": (_I, [_P, _I, _P, _P, _P]),
    "orb_extractor_set_opencv_rounding": (_I, [_P, _I]),
    "orb_context_destroy": (_V, [_P]),
    "orb_context_synchronize": (_I, [_P]),
    "orb_extractor_create": (_I, [_P, _I, _F, _I, _I, _I, _P]),
    "orb_extractor_destroy": (_V, [_P]),
    "orb_extractor_levels": (_I, [_P]),
    "orb_extractor_scale_factor": (_F, [_P]),
    "orb_extractor_scales": (_I, [_P, _P, _P, _P, _P]),
    "orb_extractor_quotas": (_I, [_P, _P]),
    "orb_extractor_capacity": (_I, [_P, _I, _I]),
    "orb_extract": (_I, [_P, _P, _I, _I, _S, _P, _I, _P, _P]),
    "orb_extract_device": (_I, [_P, _P, _I, _I, _S]),
    "orb_extractor_results": (_I, [_P, _P, _I, _P, _P]),
    "orb_extractor_pyramid_level": (_I, [_P, _I, _P, _S, _P, _P]),
    "orb_extract_color": (_I, [_P, _P, _I, _I, _S, _I, _I, _P, _I, _P, _P]),
    "orb_cvt_color_gray": (_I, [_P, _P, _I, _I, _S, _I, _I, _P, _S]),
    "orb_rectify_create": (_I, [_P, _P, _P, _I, _P, _P, _I, _I, _P]),
    "orb_rectify_from_maps": (_I, [_P, _P, _P, _I, _I, _P]),
    "orb_rectify_destroy": (_V, [_P]),
    "orb_rectify_maps": (_I, [_P, _P, _P]),
    "orb_remap": (_I, [_P, _P, _I, _I, _S, _P, _S]),
    "orb_extract_rectified": (_I, [_P, _P, _P, _I, _I, _S, _P, _I, _P, _P]),
    "orb_descriptor_distance": (_I, [_P, _P]),
    "orb_match_bf": (_I, [_P, _P, _I, _P, _I, _I, _F, _P, _P]),
    "orb_stereo_match": (_I, [_P, _P, _P, _P, _I, _P, _P, _I, _F, _F, _P, _P]),
    "orb_stereo_frame": (_I, [_P, _P, _P, _P, _I, _I, _S, _F, _F, _P, _P, _P, _P, _P, _P, _I, _P, _P]),
    "orb_context_stream": (_P, [_P]),
    "orb_hbm_copy_time": (_I, [_P, _S, _I, _P]),
    "orb_knn2": (_I, [_P, _P, _I, _P, ctypes.c_int64, _I, _F, _P, _P, _P]),
    "orb_knn2_device": (_I, [_P, _P, _I, _P, ctypes.c_int64, _I, _F, _P, _P, _P]),
    "orb_knn2_device_noalloc": (_I, [_P, _P, _I, _P, ctypes.c_int64, _I, _F, _P, _P, _P]),
    "orb_knn2_time": (_I, [_P, _P, _I, _P, ctypes.c_int64, _I, _F, _I, _P, _P]),
    "orb_batch_create": (_I, [_P, _I, _P, _P, _P, _F, _I, _I, _I, _P]),
    "orb_batch_destroy": (_V, [_P]),
    "orb_batch_size": (_I, [_P]),
    "orb_batch_image": (_I, [_P, _I, _P, _P, _P, _P]),
    "orb_batch_upload": (_I, [_P, _I, _P, _S]),
    "orb_batch_set_stereo": (_I, [_P, _I, _P, _P, _F, _F]),
    "orb_batch_run": (_I, [_P, _I]),
    "orb_batch_set_branches": (_I, [_P, _I]),
    "orb_batch_autotune": (_I, [_P, _I, _P, _P]),
    "orb_batch_synchronize": (_I, [_P]),
    "orb_batch_stage_count": (_I, [_P]),
    "orb_batch_stage_times": (_I, [_P, _I, _P]),
    "orb_batch_results": (_I, [_P, _I, _P, _I, _P, _P]),
    "orb_batch_stereo_results": (_I, [_P, _I, _P, _P, _I, _P]),
    "orb_server_create": (_I, [_P, _I, _P, _P, _P, _F, _I, _I, _I, _I, _F, _S, _I, _I, _P]),
    "orb_server_destroy": (_V, [_P]),
    "orb_server_clock": (ctypes.c_double, []),
    "orb_server_submit": (_I, [_P, _I, _P, _S, ctypes.c_double, _P]),
    "orb_server_poll": (_I, [_P, _P, _I, _P]),
    "orb_server_fetch": (_I, [_P, ctypes.c_uint64, _P, _I, _P, _P, _P, _P]),
    "orb_server_release": (_I, [_P, ctypes.c_uint64]),
    "orb_server_drain": (_I, [_P, ctypes.c_double]),
    "orb_server_stats": (_I, [_P, _P, _P, _P, _P]),
    "orb_server_replay": (_I, [_P, _I, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _P]),
    "orb_undistort_keypoints": (_I, [_P, _I, _P, _P, _P, _I, _P]),
    "orb_rgbd_depth": (_I, [_P, _I, _P, _P, _P, _I, _I, _S, _F, _P, _P]),
    "orb_rgbd_frame": (_I, [_P, _P, _I, _I, _S, _P, _S, _P, _P, _I, _F, _P, _P, _I, _P, _P, _P, _P]),
    "orb_vocabulary_load_text": (_I, [_P, ctypes.c_char_p, _P]),
    "orb_vocabulary_create": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "orb_vocabulary_destroy": (_V, [_P]),
    "orb_vocabulary_info": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "orb_vocabulary_transform": (_I, [_P, _P, _I, _I, _P, _P, _P]),
    "orb_vocabulary_bow": (_I, [_P, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P]),
    "orb_vocabulary_bow_time": (_I, [_P, _I, _P, _P, _I, _I, _P]),
    "orb_distinctive_descriptors": (_I, [_P, _I, _P, _P, _P, _P]),
    "orb_search_for_initialization": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _F, _I, _P, _P]),
    "orb_search_by_projection": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _F, _F, _P, _P, _P]),
    "orb_search_by_projection_last": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _P, _P, _F, _I, _P,
                                           _P, _P]),
    "orb_search_for_initialization_time": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _F, _I, _P, _P, _I,
                                                _P]),
    "orb_search_by_projection_time": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _F, _F, _P, _P, _P,
                                           _I, _P]),
    "orb_search_by_projection_last_time": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _P, _P, _F, _I,
                                                _P, _P, _P, _I, _P]),
    "orb_distinctive_descriptors_time": (_I, [_P, _I, _P, _P, _I, _P]),
    "orb_extractor_debug_image": (_I, [_P, _I, _I, _P, _S]),
    "orb_extractor_debug_candidates": (_I, [_P, _I, _P, _I, _P]),
}

_lib = None


class OrbError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} failed: {_ERRORS.get(rc, rc)} ({rc})")
        self.rc = rc


def lib():
    """Load liborb_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise OSError(f"{library_path} is missing: run orb_slam2_detailed_comments_amd/build.py "
                          "(or __graft_entry__.build())")
        L = ctypes.CDLL(library_path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(fn, rc):
    if rc != ORB_OK:
        raise OrbError(fn, rc)


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _view(owner, address, nbytes, dtype):
    """A numpy view of `nbytes` at `address` (memory that `owner` holds); the
    view keeps `owner` alive."""
    buf = (ctypes.c_uint8 * nbytes).from_address(address)
    buf._owner = owner
    return np.frombuffer(buf, dtype)


class Context:
    """A HIP device and the stream the front end runs on."""

    _default = None

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check("orb_context_create", lib().orb_context_create(device, ctypes.byref(h)))
        self.h = h

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def synchronize(self):
        _check("orb_context_synchronize", lib().orb_context_synchronize(self.h))

    def set_round_limit(self, rounds):
        """Parallel rounds of the windowed searches before their in-order pass
        (0..48; 0 runs the in-order pass alone)."""
        _check("orb_context_set_round_limit", lib().orb_context_set_round_limit(self.h, int(rounds)))

    def last_rounds(self):
        """(rounds, fell back to the in-order pass) of the last windowed search."""
        seq = ctypes.c_int()
        r = lib().orb_context_last_rounds(self.h, ctypes.byref(seq))
        _check("orb_context_last_rounds", min(r, 0))
        return r, bool(seq.value)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orb_context_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class ORBextractor:
    """ORBextractor(nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST).

    Mirrors include/ORBextractor.h:93-162; the pixel work runs on the GPU.
    """

    def __init__(self, nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST, context=None, opencv_sse2=True):
        """opencv_sse2: round resize and the Gaussian as x86 builds of OpenCV
        2.4/3.x do (their SSE2 kernels; the default), else as OpenCV's generic
        scalar code."""
        self.ctx = context or Context.default()
        h = ctypes.c_void_p()
        _check("orb_extractor_create", lib().orb_extractor_create(self.ctx.h, int(nfeatures), float(scaleFactor),
                                                                  int(nlevels), int(iniThFAST), int(minThFAST),
                                                                  ctypes.byref(h)))
        self.h = h
        self.nfeatures, self.nlevels = int(nfeatures), int(nlevels)
        self._size = None
        if not opencv_sse2:
            _check("orb_extractor_set_opencv_rounding", lib().orb_extractor_set_opencv_rounding(self.h, 0))

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orb_extractor_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- getters (ORBextractor.h:119-159)
    def GetLevels(self):
        return lib().orb_extractor_levels(self.h)

    def GetScaleFactor(self):
        return lib().orb_extractor_scale_factor(self.h)

    def _scales(self):
        n = self.nlevels
        out = [np.zeros(n, np.float32) for _ in range(4)]
        _check("orb_extractor_scales", lib().orb_extractor_scales(self.h, *[_ptr(a) for a in out]))
        return out

    def GetScaleFactors(self):
        return self._scales()[0]

    def GetInverseScaleFactors(self):
        return self._scales()[1]

    def GetScaleSigmaSquares(self):
        return self._scales()[2]

    def GetInverseScaleSigmaSquares(self):
        return self._scales()[3]

    def quotas(self):
        q = np.zeros(self.nlevels, np.int32)
        _check("orb_extractor_quotas", lib().orb_extractor_quotas(self.h, _ptr(q)))
        return q

    def capacity(self, width, height):
        rc = lib().orb_extractor_capacity(self.h, int(width), int(height))
        if rc < 0:
            raise OrbError("orb_extractor_capacity", rc)
        return rc

    def set_tile_heights(self, tile_h):
        """Level kernel tiles per level: 64 x 32 (32) or 64 x 16 (16)."""
        th = np.ascontiguousarray(tile_h, np.int32)
        _check("orb_extractor_set_tile_heights", lib().orb_extractor_set_tile_heights(self.h, len(th), _ptr(th)))

    # -- operator() (ORBextractor.h:111)
    def __call__(self, image, mask=None):
        """Returns (keypoints[N] KEYPOINT_DTYPE, descriptors[N, 32] uint8)."""
        img = np.asarray(image)
        if img.dtype != np.uint8 or img.ndim != 2:
            raise ValueError("image must be a 2-D uint8 (CV_8UC1) array")
        if img.size == 0:
            return np.zeros(0, KEYPOINT_DTYPE), np.zeros((0, 32), np.uint8)
        if img.strides[1] != 1:
            img = np.ascontiguousarray(img)
        h, w = img.shape
        cap = self.capacity(w, h)
        kps = np.zeros(cap, KEYPOINT_DTYPE)
        desc = np.zeros((cap, 32), np.uint8)
        n = ctypes.c_int()
        _check("orb_extract", lib().orb_extract(self.h, _ptr(img), w, h, img.strides[0], _ptr(kps), cap,
                                                _ptr(desc), ctypes.byref(n)))
        self._size = (w, h)
        return kps[:n.value].copy(), desc[:n.value].copy()

    def extract_color(self, image, rgb=True):
        """The same for an h x w x 3|4 colour image, converted on the device as
        cvtColor(CV_RGB2GRAY / CV_BGR2GRAY / ...A2GRAY) does (Tracking.cc:250-276)."""
        img = np.ascontiguousarray(image, np.uint8)
        if img.ndim != 3 or img.shape[2] not in (3, 4):
            raise ValueError("image must be h x w x 3 or 4 uint8")
        h, w, c = img.shape
        if img.size == 0:
            return np.zeros(0, KEYPOINT_DTYPE), np.zeros((0, 32), np.uint8)
        cap = self.capacity(w, h)
        kps = np.zeros(cap, KEYPOINT_DTYPE)
        desc = np.zeros((cap, 32), np.uint8)
        n = ctypes.c_int()
        _check("orb_extract_color", lib().orb_extract_color(self.h, _ptr(img), w, h, w * c, c, 1 if rgb else 0,
                                                            _ptr(kps), cap, _ptr(desc), ctypes.byref(n)))
        self._size = (w, h)
        return kps[:n.value].copy(), desc[:n.value].copy()

    def pyramid_level(self, level):
        w, h = ctypes.c_int(), ctypes.c_int()
        _check("orb_extractor_pyramid_level",
               lib().orb_extractor_pyramid_level(self.h, level, None, 0, ctypes.byref(w), ctypes.byref(h)))
        out = np.zeros((h.value, w.value), np.uint8)
        _check("orb_extractor_pyramid_level",
               lib().orb_extractor_pyramid_level(self.h, level, _ptr(out), w.value, None, None))
        return out

    @property
    def mvImagePyramid(self):
        return [self.pyramid_level(l) for l in range(self.nlevels)]

    def debug_image(self, level, kind):
        w, h = ctypes.c_int(), ctypes.c_int()
        _check("orb_extractor_pyramid_level",
               lib().orb_extractor_pyramid_level(self.h, level, None, 0, ctypes.byref(w), ctypes.byref(h)))
        out = np.zeros((h.value, w.value), np.uint8)
        _check("orb_extractor_debug_image", lib().orb_extractor_debug_image(self.h, level, kind, _ptr(out), w.value))
        return out

    def debug_candidates(self, level):
        n = ctypes.c_int()
        rc = lib().orb_extractor_debug_candidates(self.h, level, None, 0, ctypes.byref(n))
        if rc not in (ORB_OK, ORB_E_CAPACITY):
            raise OrbError("orb_extractor_debug_candidates", rc)
        buf = np.zeros((max(n.value, 1), 3), np.int32)
        _check("orb_extractor_debug_candidates",
               lib().orb_extractor_debug_candidates(self.h, level, _ptr(buf), n.value, ctypes.byref(n)))
        return buf[:n.value].copy()


class ORBmatcher:
    """ORBmatcher's thresholds (src/ORBmatcher.cc:49-51) and distance."""
    TH_HIGH = 100
    TH_LOW = 50
    HISTO_LENGTH = 30

    @staticmethod
    def DescriptorDistance(a, b):
        a = np.ascontiguousarray(a, np.uint8).reshape(-1)
        b = np.ascontiguousarray(b, np.uint8).reshape(-1)
        if a.size < 32 or b.size < 32:
            raise ValueError("descriptors are 32-byte rows")
        return lib().orb_descriptor_distance(_ptr(a), _ptr(b))


def match_bf(query, train, th_low=ORBmatcher.TH_LOW, ratio=0.6, context=None):
    """Brute-force best/second-best matching on the GPU: (idx[-1 if none], dist)."""
    ctx = context or Context.default()
    q = np.ascontiguousarray(query, np.uint8)
    t = np.ascontiguousarray(train, np.uint8)
    idx = np.zeros(len(q), np.int32)
    dist = np.zeros(len(q), np.int32)
    _check("orb_match_bf", lib().orb_match_bf(ctx.h, _ptr(q), len(q), _ptr(t), len(t), int(th_low), float(ratio),
                                             _ptr(idx), _ptr(dist)))
    return idx, dist


def knn2(query, map_rows, ratio=0.7, method="mfma", context=None):
    """Dense k=2 Hamming search: (idx[nq, 2], dist[nq, 2], accept[nq])."""
    ctx = context or Context.default()
    q = np.ascontiguousarray(query, np.uint8)
    m = np.ascontiguousarray(map_rows, np.uint8)
    idx = np.zeros((len(q), 2), np.int32)
    dist = np.zeros((len(q), 2), np.int32)
    acc = np.zeros(len(q), np.uint8)
    meth = {"popcount": 0, "mfma": 1}[method]
    _check("orb_knn2", lib().orb_knn2(ctx.h, _ptr(q), len(q), _ptr(m), len(m), meth, float(ratio), _ptr(idx),
                                      _ptr(dist), _ptr(acc)))
    return idx, dist, acc.astype(bool)


def compute_stereo_matches(left, right, kl, dl, kr, dr, bf, b):
    """Frame::ComputeStereoMatches on the GPU -> (mvuRight, mvDepth)."""
    kl = np.ascontiguousarray(kl, KEYPOINT_DTYPE)
    kr = np.ascontiguousarray(kr, KEYPOINT_DTYPE)
    dl = np.ascontiguousarray(dl, np.uint8)
    dr = np.ascontiguousarray(dr, np.uint8)
    uR = np.zeros(len(kl), np.float32)
    dep = np.zeros(len(kl), np.float32)
    _check("orb_stereo_match", lib().orb_stereo_match(left.h, right.h, _ptr(kl), _ptr(dl), len(kl), _ptr(kr),
                                                     _ptr(dr), len(kr), float(bf), float(b), _ptr(uR),
                                                     _ptr(dep)))
    return uR, dep


def stereo_frame(left, right, im_left, im_right, bf, b):
    """Both extractions plus stereo matching in one device pass.

    Returns (kl, dl, kr, dr, mvuRight, mvDepth)."""
    imL = np.ascontiguousarray(im_left, np.uint8)
    imR = np.ascontiguousarray(im_right, np.uint8)
    if imL.shape != imR.shape:
        raise ValueError("left and right images must have the same size")
    h, w = imL.shape
    cap = max(left.capacity(w, h), right.capacity(w, h))
    kl = np.zeros(cap, KEYPOINT_DTYPE)
    kr = np.zeros(cap, KEYPOINT_DTYPE)
    dl = np.zeros((cap, 32), np.uint8)
    dr = np.zeros((cap, 32), np.uint8)
    uR = np.zeros(cap, np.float32)
    dep = np.zeros(cap, np.float32)
    nl, nr = ctypes.c_int(), ctypes.c_int()
    _check("orb_stereo_frame", lib().orb_stereo_frame(left.h, right.h, _ptr(imL), _ptr(imR), w, h, imL.strides[0],
                                                     float(bf), float(b), _ptr(kl), _ptr(dl), ctypes.byref(nl),
                                                     _ptr(kr), _ptr(dr), ctypes.byref(nr), cap, _ptr(uR),
                                                     _ptr(dep)))
    n = nl.value
    return kl[:n].copy(), dl[:n].copy(), kr[:nr.value].copy(), dr[:nr.value].copy(), uR[:n].copy(), dep[:n].copy()


def rgbd_frame(extractor, image, depth, K, D, bf):
    """The RGB-D Frame constructor's hot path in one device pass: extraction,
    UndistortKeyPoints and ComputeStereoFromRGBD (depth None: the monocular
    constructor's, extraction and UndistortKeyPoints; mvuRight and mvDepth
    are then all -1).

    Returns (keypoints, descriptors, keypoints_un, mvuRight, mvDepth)."""
    img = np.ascontiguousarray(image, np.uint8)
    h, w = img.shape
    dep_img = None
    if depth is not None:
        dep_img = np.asarray(depth, np.float32)
        if dep_img.ndim != 2 or dep_img.strides[1] != 4 or dep_img.strides[0] < 4 * dep_img.shape[1]:
            dep_img = np.ascontiguousarray(dep_img)
        if dep_img.shape != (h, w):
            raise ValueError("depth must have the image's size")
    K = np.ascontiguousarray(K, np.float32)
    D = np.ascontiguousarray(D, np.float32)
    cap = extractor.capacity(w, h)
    kps = np.zeros(cap, KEYPOINT_DTYPE)
    kun = np.zeros(cap, KEYPOINT_DTYPE)
    desc = np.zeros((cap, 32), np.uint8)
    uR = np.full(cap, -1, np.float32)
    dep = np.full(cap, -1, np.float32)
    n = ctypes.c_int()
    _check("orb_rgbd_frame", lib().orb_rgbd_frame(extractor.h, _ptr(img), w, h, img.strides[0], _ptr(dep_img),
                                                  dep_img.strides[0] if dep_img is not None else 0, _ptr(K), _ptr(D),
                                                  len(D), float(bf), _ptr(kps),
                                                  _ptr(kun), cap, _ptr(desc), _ptr(uR), _ptr(dep), ctypes.byref(n)))
    k = n.value
    return kps[:k].copy(), desc[:k].copy(), kun[:k].copy(), uR[:k].copy(), dep[:k].copy()


class Batch:
    """A fixed set of image slots extracted together (orb_batch_* in the C ABI).

    run(mode): 0 = one launch per stage over all slots, 1 = the same captured
    in a hipGraph and replayed, 2 = per-image launches of the same kernels.
    """
    STAGE_LAUNCHES, GRAPH, PER_IMAGE, STREAMS = 0, 1, 2, 3

    def __init__(self, sizes, nfeatures, scaleFactor=1.2, nlevels=8, iniThFAST=20, minThFAST=7, context=None):
        self.ctx = context or Context.default()
        n = len(sizes)
        w = np.array([s[0] for s in sizes], np.int32)
        h = np.array([s[1] for s in sizes], np.int32)
        nf = np.broadcast_to(np.asarray(nfeatures, np.int32), (n,)).copy()
        hb = ctypes.c_void_p()
        _check("orb_batch_create", lib().orb_batch_create(self.ctx.h, n, _ptr(w), _ptr(h), _ptr(nf),
                                                          float(scaleFactor), int(nlevels), int(iniThFAST),
                                                          int(minThFAST), ctypes.byref(hb)))
        self.h = hb
        self.sizes = [tuple(s) for s in sizes]
        self.nfeatures = nf
        self.nlevels = int(nlevels)
        self.npairs = 0

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __len__(self):
        return len(self.sizes)

    def upload(self, slot, image):
        img = np.ascontiguousarray(image, np.uint8)
        if img.shape != (self.sizes[slot][1], self.sizes[slot][0]):
            raise ValueError(f"slot {slot} expects {self.sizes[slot]}, got {img.shape[::-1]}")
        _check("orb_batch_upload", lib().orb_batch_upload(self.h, slot, _ptr(img), img.strides[0]))

    def upload_all(self, images, slots=None):
        """Enqueue host images into slots (default 0..n-1) without waiting."""
        slots = list(range(len(images))) if slots is None else list(slots)
        imgs = [np.ascontiguousarray(im, np.uint8) for im in images]
        for s, im in zip(slots, imgs):
            if im.shape != (self.sizes[s][1], self.sizes[s][0]):
                raise ValueError(f"slot {s} expects {self.sizes[s]}, got {im.shape[::-1]}")
        sl = np.asarray(slots, np.int32)
        ptrs = (ctypes.c_void_p * len(imgs))(*[im.ctypes.data for im in imgs])
        strides = np.asarray([im.strides[0] for im in imgs], np.uint64)
        _check("orb_batch_upload_images", lib().orb_batch_upload_images(self.h, len(imgs), _ptr(sl), ptrs,
                                                                        _ptr(strides)))

    def read_all(self, capacity=None, out=None):
        """Every slot's (keypoints, descriptors) and every pair's (uRight,
        depth) after one wait; `out` (from a previous call) is refilled."""
        cap = int(capacity or (int(self.nfeatures.max()) + 64 * self.nlevels))
        n, npairs = len(self.sizes), self.npairs
        if out is None:
            out = {"kps": np.zeros((n, cap), KEYPOINT_DTYPE), "desc": np.zeros((n, cap, 32), np.uint8),
                   "counts": np.zeros(n, np.int32), "uR": np.zeros((max(npairs, 1), cap), np.float32),
                   "depth": np.zeros((max(npairs, 1), cap), np.float32)}
        _check("orb_batch_read_all", lib().orb_batch_read_all(self.h, _ptr(out["kps"]), _ptr(out["desc"]), cap,
                                                              _ptr(out["counts"]), _ptr(out["uR"]),
                                                              _ptr(out["depth"])))
        return out

    def pinned_images(self):
        """The batch's pinned host image buffers, one (height, width) uint8
        view per slot (rows at the device pitch).  Write frames into them,
        then upload_pinned().  The views keep the batch alive."""
        n = len(self.sizes)
        ptrs = (ctypes.c_void_p * n)()
        pitches = np.zeros(n, np.uint64)
        _check("orb_batch_pinned_images", lib().orb_batch_pinned_images(self.h, ptrs, _ptr(pitches)))
        out = []
        for s in range(n):
            w, h = self.sizes[s]
            full = _view(self, ptrs[s], int(pitches[s]) * h, np.uint8).reshape(h, int(pitches[s]))
            out.append(full[:, :w])
        return out

    def upload_pinned(self):
        """Enqueue the pinned images' transfer; the next run() follows it."""
        _check("orb_batch_upload_pinned", lib().orb_batch_upload_pinned(self.h))

    def read_async(self):
        """Enqueue the transfer of every result into pinned host memory."""
        _check("orb_batch_read_async", lib().orb_batch_read_async(self.h))

    def pinned_results(self, slot):
        """(keypoints, descriptors) of the last read_async(), after waiting for
        it: views into pinned memory, valid until the next read_async()."""
        kp, de, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
        _check("orb_batch_pinned_results", lib().orb_batch_pinned_results(self.h, slot, ctypes.byref(kp),
                                                                          ctypes.byref(de), ctypes.byref(n)))
        k = n.value
        if k == 0:
            return np.zeros(0, KEYPOINT_DTYPE), np.zeros((0, 32), np.uint8)
        kps = _view(self, kp.value, k * KEYPOINT_DTYPE.itemsize, KEYPOINT_DTYPE)
        desc = _view(self, de.value, k * 32, np.uint8).reshape(k, 32)
        return kps, desc

    def pinned_stereo(self, pair):
        """(uRight, depth) of pair `pair` from the last read_async(), as views."""
        ur, dp, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
        _check("orb_batch_pinned_stereo", lib().orb_batch_pinned_stereo(self.h, pair, ctypes.byref(ur),
                                                                        ctypes.byref(dp), ctypes.byref(n)))
        k = n.value
        if k == 0:
            return np.zeros(0, np.float32), np.zeros(0, np.float32)
        return _view(self, ur.value, 4 * k, np.float32), _view(self, dp.value, 4 * k, np.float32)

    def set_stereo(self, left_slots, right_slots, bf, baseline):
        l = np.ascontiguousarray(left_slots, np.int32)
        r = np.ascontiguousarray(right_slots, np.int32)
        _check("orb_batch_set_stereo", lib().orb_batch_set_stereo(self.h, len(l), _ptr(l), _ptr(r), float(bf),
                                                                  float(baseline)))
        self.npairs = len(l)

    def set_branches(self, branches):
        """Split the mode-1 graph into this many independent branches."""
        _check("orb_batch_set_branches", lib().orb_batch_set_branches(self.h, int(branches)))

    def autotune(self, steps=10):
        """Pick the fastest graph branch count (1, 2, 3, 4, 6 or 8) by timing
        `steps` replays of each; returns (best, {branches: ms per replay})."""
        best = ctypes.c_int(0)
        ms = np.zeros(8, np.float32)
        _check("orb_batch_autotune", lib().orb_batch_autotune(self.h, int(steps), ctypes.byref(best), _ptr(ms)))
        return best.value, {k + 1: float(v) for k, v in enumerate(ms) if v >= 0}

    def set_tile_heights(self, tile_h):
        """Level kernel tiles per level (32 or 16) for every slot."""
        th = np.ascontiguousarray(tile_h, np.int32)
        _check("orb_batch_set_tile_heights", lib().orb_batch_set_tile_heights(self.h, len(th), _ptr(th)))

    def autotune_tiles(self, steps=10):
        """Per level, the faster of 64 x 32 and 64 x 16 tiles on the current
        images; returns (tile heights, {level: ms at 32}, {level: ms at 16})."""
        th = np.zeros(self.nlevels, np.int32)
        m32 = np.zeros(self.nlevels, np.float32)
        m16 = np.zeros(self.nlevels, np.float32)
        _check("orb_batch_autotune_tiles", lib().orb_batch_autotune_tiles(self.h, int(steps), _ptr(th), _ptr(m32),
                                                                          _ptr(m16)))
        return th.tolist(), dict(enumerate(m32.tolist())), dict(enumerate(m16.tolist()))

    def run(self, mode=GRAPH):
        _check("orb_batch_run", lib().orb_batch_run(self.h, int(mode)))

    def synchronize(self):
        _check("orb_batch_synchronize", lib().orb_batch_synchronize(self.h))

    def stage_times(self, steps):
        n = lib().orb_batch_stage_count(self.h)
        ms = np.zeros(n, np.float32)
        _check("orb_batch_stage_times", lib().orb_batch_stage_times(self.h, int(steps), _ptr(ms)))
        names = [f"pyramid_l{l}" for l in range(n - 4)] + ["octree", "describe", "stereo_match", "stereo_median"]
        return dict(zip(names, ms.tolist()))

    def results(self, slot):
        w, h = self.sizes[slot]
        cap = int(self.nfeatures[slot]) + 64 * self.nlevels
        kps = np.zeros(cap, KEYPOINT_DTYPE)
        desc = np.zeros((cap, 32), np.uint8)
        n = ctypes.c_int()
        _check("orb_batch_results", lib().orb_batch_results(self.h, slot, _ptr(kps), cap, _ptr(desc),
                                                            ctypes.byref(n)))
        return kps[:n.value].copy(), desc[:n.value].copy()

    def stereo_results(self, pair):
        cap = 1 << 16
        uR = np.zeros(cap, np.float32)
        dep = np.zeros(cap, np.float32)
        n = ctypes.c_int()
        _check("orb_batch_stereo_results", lib().orb_batch_stereo_results(self.h, pair, _ptr(uR), _ptr(dep), cap,
                                                                          ctypes.byref(n)))
        return uR[:n.value].copy(), dep[:n.value].copy()


class _ServeInfo(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("stream", ctypes.c_int), ("status", ctypes.c_int),
                ("n_keypoints", ctypes.c_int), ("n_matches", ctypes.c_int), ("t_arrival", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


class _ReplayStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("refused", ctypes.c_int64), ("errors", ctypes.c_int64),
                ("frames_per_s", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double)]


class Server:
    """Continuous-batching serving loop for camera streams (orb_server_* in the C ABI).

    streams: list of (width, height, nfeatures).  Each submitted frame is
    extracted and matched (best/second, th_low, ratio) against its stream's
    previous frame, whose descriptors stay resident in HBM.
    """

    def __init__(self, streams, scaleFactor=1.2, nlevels=8, iniThFAST=20, minThFAST=7, th_low=50, ratio=0.6,
                 hbm_budget=2 << 30, max_inflight=3, max_batch=64, context=None):
        self.ctx = context or Context.default()
        n = len(streams)
        w = np.array([s[0] for s in streams], np.int32)
        h = np.array([s[1] for s in streams], np.int32)
        nf = np.array([s[2] for s in streams], np.int32)
        hs = ctypes.c_void_p()
        _check("orb_server_create", lib().orb_server_create(self.ctx.h, n, _ptr(w), _ptr(h), _ptr(nf),
                                                            float(scaleFactor), int(nlevels), int(iniThFAST),
                                                            int(minThFAST), int(th_low), float(ratio),
                                                            int(hbm_budget), int(max_inflight), int(max_batch),
                                                            ctypes.byref(hs)))
        self.h = hs
        self.streams = [tuple(s) for s in streams]
        self._info = (_ServeInfo * 1024)()

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orb_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def clock():
        return lib().orb_server_clock()

    def submit(self, stream, image, t_arrival=None):
        img = np.ascontiguousarray(image, np.uint8)
        w, h, _ = self.streams[stream]
        if img.shape != (h, w):
            raise ValueError(f"stream {stream} expects {w}x{h}")
        t = ctypes.c_uint64()
        ta = self.clock() if t_arrival is None else float(t_arrival)
        _check("orb_server_submit", lib().orb_server_submit(self.h, stream, _ptr(img), img.strides[0], ta,
                                                            ctypes.byref(t)))
        return t.value

    def poll(self):
        n = ctypes.c_int()
        _check("orb_server_poll", lib().orb_server_poll(self.h, self._info, len(self._info), ctypes.byref(n)))
        return [dict(ticket=i.ticket, stream=i.stream, status=i.status, n_keypoints=i.n_keypoints,
                     n_matches=i.n_matches, t_arrival=i.t_arrival, t_admit=i.t_admit, t_done=i.t_done)
                for i in self._info[:n.value]]

    def fetch(self, ticket, capacity=8192):
        kps = np.zeros(capacity, KEYPOINT_DTYPE)
        desc = np.zeros((capacity, 32), np.uint8)
        idx = np.zeros(capacity, np.int32)
        dist = np.zeros(capacity, np.int32)
        n = ctypes.c_int()
        _check("orb_server_fetch", lib().orb_server_fetch(self.h, ticket, _ptr(kps), capacity, _ptr(desc), _ptr(idx),
                                                          _ptr(dist), ctypes.byref(n)))
        k = n.value
        return kps[:k].copy(), desc[:k].copy(), idx[:k].copy(), dist[:k].copy()

    def release(self, ticket):
        _check("orb_server_release", lib().orb_server_release(self.h, ticket))

    def drain(self, timeout=60.0):
        _check("orb_server_drain", lib().orb_server_drain(self.h, float(timeout)))

    def replay(self, frames, fps, duration, seed=1):
        """Open-loop replay from native threads (orb_server_replay): frames[s]
        is stream s's sequence, replayed forwards then backwards at Poisson
        arrivals of `fps` per stream for `duration` seconds.  Returns the
        request count, refusals, errors, frames/s and p50/p99/max latency."""
        n = len(frames)
        stacks = []
        for s, seq in enumerate(frames):
            w, h, _ = self.streams[s]
            a = np.ascontiguousarray(np.stack([np.asarray(f, np.uint8) for f in seq]))
            if a.shape[1:] != (h, w):
                raise ValueError(f"stream {s} expects {w}x{h}")
            stacks.append(a)
        ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in stacks])
        nf = np.array([len(a) for a in stacks], np.int32)
        st = _ReplayStats()
        _check("orb_server_replay", lib().orb_server_replay(self.h, n, ptrs, _ptr(nf), float(fps), float(duration),
                                                            ctypes.c_uint64(seed), ctypes.byref(st)))
        return {"requests": st.requests, "refused": st.refused, "errors": st.errors,
                "frames_per_s": round(st.frames_per_s, 1), "p50_ms": round(st.p50_ms, 3),
                "p99_ms": round(st.p99_ms, 3), "max_ms": round(st.max_ms, 3)}

    def stats(self):
        b, r, u, sl = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_size_t(), ctypes.c_int()
        _check("orb_server_stats", lib().orb_server_stats(self.h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(u),
                                                          ctypes.byref(sl)))
        return dict(batches=b.value, requests=r.value, hbm_used=u.value, slots=sl.value)


def undistort_keypoints(kps, K, D, context=None):
    """Frame::UndistortKeyPoints: keypoints (KEYPOINT_DTYPE) with x, y put
    through cv::undistortPoints(K, D, R = I, P = K); K = (fx, fy, cx, cy),
    D = 4 or 5 coefficients k1 k2 p1 p2 [k3]."""
    ctx = context or Context.default()
    kps = np.ascontiguousarray(kps, KEYPOINT_DTYPE)
    K = np.ascontiguousarray(K, np.float32)
    D = np.ascontiguousarray(D, np.float32)
    if K.shape != (4,):
        raise ValueError("K is (fx, fy, cx, cy)")
    out = np.empty_like(kps)
    _check("orb_undistort_keypoints", lib().orb_undistort_keypoints(ctx.h, len(kps), _ptr(kps), _ptr(K), _ptr(D),
                                                                    len(D), _ptr(out)))
    return out


def rgbd_depth(kps, kps_un, depth, bf, context=None):
    """Frame::ComputeStereoFromRGBD: (uRight, depth) per keypoint from a
    float32 depth image."""
    ctx = context or Context.default()
    kps = np.ascontiguousarray(kps, KEYPOINT_DTYPE)
    kun = np.ascontiguousarray(kps_un, KEYPOINT_DTYPE)
    if len(kun) != len(kps):
        raise ValueError("kps and kps_un differ in length")
    depth = np.asarray(depth, np.float32)
    if depth.ndim != 2 or depth.strides[1] != 4 or depth.strides[0] < 4 * depth.shape[1]:
        depth = np.ascontiguousarray(depth)       # rows may be strided, pixels not
    uR = np.empty(len(kps), np.float32)
    dep = np.empty(len(kps), np.float32)
    _check("orb_rgbd_depth", lib().orb_rgbd_depth(ctx.h, len(kps), _ptr(kps), _ptr(kun), _ptr(depth), depth.shape[1],
                                                  depth.shape[0], depth.strides[0], float(bf), _ptr(uR), _ptr(dep)))
    return uR, dep


def cvt_color_gray(image, rgb=True, context=None):
    """cvtColor to gray of an h x w x 3|4 uint8 image on the device."""
    ctx = context or Context.default()
    img = np.ascontiguousarray(image, np.uint8)
    h, w, c = img.shape
    out = np.zeros((h, w), np.uint8)
    _check("orb_cvt_color_gray", lib().orb_cvt_color_gray(ctx.h, _ptr(img), w, h, w * c, c, 1 if rgb else 0,
                                                          _ptr(out), w))
    return out


class Rectifier:
    """Stereo rectification maps in HBM (cv::initUndistortRectifyMap, CV_32FC1)
    and cv::remap(INTER_LINEAR) on the device (stereo_euroc.cc:124-188).
    Rectifier(K, D, R, P, width, height) or Rectifier(mapx=..., mapy=...)."""

    def __init__(self, K=None, D=None, R=None, P=None, width=None, height=None, mapx=None, mapy=None, context=None):
        self.ctx = context or Context.default()
        h = ctypes.c_void_p()
        if mapx is not None:
            mx = np.ascontiguousarray(mapx, np.float32)
            my = np.ascontiguousarray(mapy, np.float32)
            self.height, self.width = mx.shape
            _check("orb_rectify_from_maps", lib().orb_rectify_from_maps(self.ctx.h, _ptr(mx), _ptr(my), self.width,
                                                                        self.height, ctypes.byref(h)))
        else:
            K = np.ascontiguousarray(K, np.float64).reshape(9)
            D = np.ascontiguousarray(D if D is not None else [], np.float64).reshape(-1)
            R = None if R is None else np.ascontiguousarray(R, np.float64).reshape(9)
            P = np.ascontiguousarray(np.asarray(P, np.float64).reshape(3, -1)[:, :3]).reshape(9)
            self.width, self.height = int(width), int(height)
            _check("orb_rectify_create",
                   lib().orb_rectify_create(self.ctx.h, _ptr(K), _ptr(D) if len(D) else None, len(D), _ptr(R),
                                            _ptr(P), self.width, self.height, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().orb_rectify_destroy(self.h)
            self.h = None

    __del__ = close

    def maps(self):
        mx = np.zeros((self.height, self.width), np.float32)
        my = np.zeros((self.height, self.width), np.float32)
        _check("orb_rectify_maps", lib().orb_rectify_maps(self.h, _ptr(mx), _ptr(my)))
        return mx, my

    def remap(self, image):
        img = np.ascontiguousarray(image, np.uint8)
        out = np.zeros((self.height, self.width), np.uint8)
        _check("orb_remap", lib().orb_remap(self.h, _ptr(img), img.shape[1], img.shape[0], img.shape[1], _ptr(out),
                                            self.width))
        return out

    def extract(self, extractor, image):
        """Rectify on the device, then extract: (keypoints, descriptors)."""
        img = np.ascontiguousarray(image, np.uint8)
        cap = extractor.capacity(self.width, self.height)
        kps = np.zeros(cap, KEYPOINT_DTYPE)
        desc = np.zeros((cap, 32), np.uint8)
        n = ctypes.c_int()
        _check("orb_extract_rectified",
               lib().orb_extract_rectified(extractor.h, self.h, _ptr(img), img.shape[1], img.shape[0], img.shape[1],
                                           _ptr(kps), cap, _ptr(desc), ctypes.byref(n)))
        return kps[:n.value].copy(), desc[:n.value].copy()


def _sfi_arrays(pairs):
    """Concatenates per-pair (kp1, desc1, kp2, desc2, bounds, prev) tuples
    into the flat arrays of the C ABI."""
    k1 = [np.ascontiguousarray(p[0], KEYPOINT_DTYPE) for p in pairs]
    k2 = [np.ascontiguousarray(p[2], KEYPOINT_DTYPE) for p in pairs]
    o1 = np.zeros(len(pairs) + 1, np.int32)
    o1[1:] = np.cumsum([len(k) for k in k1])
    o2 = np.zeros(len(pairs) + 1, np.int32)
    o2[1:] = np.cumsum([len(k) for k in k2])
    cat = lambda xs, dt, shape: np.concatenate(xs) if sum(len(x) for x in xs) else np.zeros(shape, dt)
    return dict(K1=cat(k1, KEYPOINT_DTYPE, 1),
                D1=cat([np.ascontiguousarray(p[1], np.uint8).reshape(-1, 32) for p in pairs], np.uint8, (1, 32)),
                K2=cat(k2, KEYPOINT_DTYPE, 1),
                D2=cat([np.ascontiguousarray(p[3], np.uint8).reshape(-1, 32) for p in pairs], np.uint8, (1, 32)),
                o1=o1, o2=o2, bounds=np.ascontiguousarray([p[4] for p in pairs], np.float32).reshape(-1),
                prev=cat([np.asarray(p[5], np.float32).reshape(-1, 2) for p in pairs], np.float32, (1, 2)).copy(),
                m12=np.zeros(max(int(o1[-1]), 1), np.int32), nm=np.zeros(len(pairs), np.int32))


def _sfi_call(ctx, a, window, nnratio, check_orientation, timing_steps=0):
    """orb_search_for_initialization on flat arrays (a from _sfi_arrays;
    a["prev"], a["m12"], a["nm"] receive the outputs).  Returns the timed
    form's ms, or None."""
    args = (ctx.h, len(a["nm"]), _ptr(a["K1"]), _ptr(a["D1"]), _ptr(a["o1"]), _ptr(a["K2"]), _ptr(a["D2"]),
            _ptr(a["o2"]), _ptr(a["bounds"]), _ptr(a["prev"]), int(window), float(nnratio),
            int(bool(check_orientation)), _ptr(a["m12"]), _ptr(a["nm"]))
    if timing_steps > 0:
        ms = ctypes.c_float()
        _check("orb_search_for_initialization_time",
               lib().orb_search_for_initialization_time(*args, int(timing_steps), ctypes.byref(ms)))
        return ms.value
    _check("orb_search_for_initialization", lib().orb_search_for_initialization(*args))
    return None


def search_for_initialization(pairs, window=100, nnratio=0.9, check_orientation=True, context=None,
                              timing_steps=0):
    """ORBmatcher::SearchForInitialization for a list of frame pairs
    (kp1, desc1, kp2, desc2, bounds=(minX, maxX, minY, maxY), prev (N1 x 2)).
    Returns per pair (nmatches, matches12, updated prev); with timing_steps
    > 0, (that list, mean ms of the device work with the inputs in HBM)."""
    ctx = context or Context.default()
    a = _sfi_arrays(pairs)
    ms = _sfi_call(ctx, a, window, nnratio, check_orientation, timing_steps)
    o1, m12, prev, nm = a["o1"], a["m12"], a["prev"], a["nm"]
    res = [(int(nm[i]), m12[o1[i]:o1[i + 1]].copy(), prev[o1[i]:o1[i + 1]].copy()) for i in range(len(pairs))]
    return (res, ms) if timing_steps > 0 else res


PROJ_POINT_DTYPE = np.dtype([("proj_x", "<f4"), ("proj_y", "<f4"), ("proj_xr", "<f4"), ("view_cos", "<f4"),
                             ("level", "<i4"), ("flags", "<i4")])
LAST_POINT_DTYPE = np.dtype([("xc", "<f4"), ("yc", "<f4"), ("zc", "<f4"), ("angle", "<f4"), ("octave", "<i4"),
                             ("flags", "<i4")])


def _projection_frames(frames, pdtype):
    """Concatenates per-frame (kp, desc, uright, occupied, bounds, points,
    point_desc, ...) tuples into the flat arrays of the C ABI."""
    cat = lambda xs, dt, shape: np.concatenate(xs) if sum(len(x) for x in xs) else np.zeros(shape, dt)
    kp = [np.ascontiguousarray(f[0], KEYPOINT_DTYPE) for f in frames]
    ko = np.zeros(len(frames) + 1, np.int32)
    ko[1:] = np.cumsum([len(k) for k in kp])
    pts = [np.ascontiguousarray(f[5], pdtype) for f in frames]
    po = np.zeros(len(frames) + 1, np.int32)
    po[1:] = np.cumsum([len(p) for p in pts])
    arrays = dict(
        kp=cat(kp, KEYPOINT_DTYPE, 1),
        desc=cat([np.ascontiguousarray(f[1], np.uint8).reshape(-1, 32) for f in frames], np.uint8, (1, 32)),
        uright=cat([np.ascontiguousarray(f[2], np.float32) for f in frames], np.float32, 1),
        occupied=cat([np.ascontiguousarray(f[3], np.uint8) for f in frames], np.uint8, 1),
        bounds=np.ascontiguousarray([f[4] for f in frames], np.float32).reshape(-1),
        pts=cat(pts, pdtype, 1),
        pdesc=cat([np.ascontiguousarray(f[6], np.uint8).reshape(-1, 32) for f in frames], np.uint8, (1, 32)),
        ko=ko, po=po)
    arrays["pm"] = np.zeros(max(int(po[-1]), 1), np.int32)
    arrays["kpm"] = np.zeros(max(int(ko[-1]), 1), np.int32)
    arrays["nm"] = np.zeros(len(frames), np.int32)
    return arrays


def _projection_results(a, n):
    return [(int(a["nm"][i]), a["pm"][a["po"][i]:a["po"][i + 1]].copy(), a["kpm"][a["ko"][i]:a["ko"][i + 1]].copy())
            for i in range(n)]


def search_by_projection(frames, scale_factors, th=3.0, nn_ratio=0.8, context=None, timing_steps=0):
    """ORBmatcher::SearchByProjection(F, vpMapPoints, th) for a list of frames
    (kp, desc, uright, occupied, bounds, points PROJ_POINT_DTYPE, point_desc).
    Returns per frame (nmatches, point_match, key_point); with timing_steps
    > 0, (that list, mean ms of the device work with the inputs in HBM)."""
    ctx = context or Context.default()
    a = _projection_frames(frames, PROJ_POINT_DTYPE)
    sc = np.ascontiguousarray(scale_factors, np.float32)
    args = (ctx.h, len(frames), _ptr(a["kp"]), _ptr(a["desc"]), _ptr(a["uright"]), _ptr(a["occupied"]), _ptr(a["ko"]),
            _ptr(a["bounds"]), _ptr(sc), len(sc), _ptr(a["pts"]), _ptr(a["pdesc"]), _ptr(a["po"]), float(th),
            float(nn_ratio), _ptr(a["pm"]), _ptr(a["kpm"]), _ptr(a["nm"]))
    ms = ctypes.c_float()
    if timing_steps > 0:
        _check("orb_search_by_projection_time",
               lib().orb_search_by_projection_time(*args, int(timing_steps), ctypes.byref(ms)))
    else:
        _check("orb_search_by_projection", lib().orb_search_by_projection(*args))
    res = _projection_results(a, len(frames))
    return (res, ms.value) if timing_steps > 0 else res


def search_by_projection_last(frames, scale_factors, th, mono=False, check_orientation=True, context=None,
                              timing_steps=0):
    """ORBmatcher::SearchByProjection(Current, Last, th, bMono) for a list of
    frames (kp, desc, uright, occupied, bounds, points LAST_POINT_DTYPE,
    point_desc, camera = (fx, fy, cx, cy, mbf, mb, tlc_z)).  Returns per
    frame (nmatches, point_match, key_point); with timing_steps > 0, (that
    list, mean ms of the device work with the inputs in HBM)."""
    ctx = context or Context.default()
    a = _projection_frames(frames, LAST_POINT_DTYPE)
    sc = np.ascontiguousarray(scale_factors, np.float32)
    cam = np.ascontiguousarray([f[7] for f in frames], np.float32).reshape(-1)
    args = (ctx.h, len(frames), _ptr(a["kp"]), _ptr(a["desc"]), _ptr(a["uright"]), _ptr(a["occupied"]), _ptr(a["ko"]),
            _ptr(a["bounds"]), _ptr(sc), len(sc), _ptr(cam), int(bool(mono)), _ptr(a["pts"]), _ptr(a["pdesc"]),
            _ptr(a["po"]), float(th), int(bool(check_orientation)), _ptr(a["pm"]), _ptr(a["kpm"]), _ptr(a["nm"]))
    ms = ctypes.c_float()
    if timing_steps > 0:
        _check("orb_search_by_projection_last_time",
               lib().orb_search_by_projection_last_time(*args, int(timing_steps), ctypes.byref(ms)))
    else:
        _check("orb_search_by_projection_last", lib().orb_search_by_projection_last(*args))
    res = _projection_results(a, len(frames))
    return (res, ms.value) if timing_steps > 0 else res


def distinctive_descriptors(points, context=None):
    """MapPoint::ComputeDistinctiveDescriptors for a list of N_p x 32
    observation-descriptor arrays: (chosen row per point, -1 if empty;
    the chosen descriptors, n_points x 32)."""
    ctx = context or Context.default()
    mats = [np.ascontiguousarray(d, np.uint8).reshape(-1, 32) for d in points]
    offs = np.zeros(len(mats) + 1, np.int32)
    offs[1:] = np.cumsum([len(d) for d in mats])
    desc = np.concatenate(mats) if offs[-1] else np.zeros((1, 32), np.uint8)
    best = np.zeros(max(len(mats), 1), np.int32)
    out = np.zeros((max(len(mats), 1), 32), np.uint8)
    _check("orb_distinctive_descriptors",
           lib().orb_distinctive_descriptors(ctx.h, len(mats), _ptr(desc), _ptr(offs), _ptr(best), _ptr(out)))
    return best[:len(mats)], out[:len(mats)]


class Vocabulary:
    """DBoW2 vocabulary tree in HBM (TemplatedVocabulary,
    Thirdparty/DBoW2/DBoW2/TemplatedVocabulary.h).

    ``Vocabulary(path=...)`` reads the ORBvoc.txt text format
    (loadFromTextFile, :1368); ``Vocabulary(k, L, scoring, weighting, parent,
    is_word, desc, weight)`` builds the same tree from file-order arrays
    (node 0 the root, parent[i] < i).
    """

    TF_IDF, TF, IDF, BINARY = 0, 1, 2, 3
    L1_NORM, L2_NORM, CHI_SQUARE, KL, BHATTACHARYYA, DOT_PRODUCT = 0, 1, 2, 3, 4, 5

    def __init__(self, k=None, L=None, scoring=0, weighting=0, parent=None, is_word=None, desc=None, weight=None,
                 path=None, context=None):
        self.ctx = context or Context.default()
        h = ctypes.c_void_p()
        if path is not None:
            _check("orb_vocabulary_load_text", lib().orb_vocabulary_load_text(self.ctx.h, os.fsencode(path),
                                                                              ctypes.byref(h)))
        else:
            parent = np.ascontiguousarray(parent, np.int32)
            n = len(parent)
            is_word = np.ascontiguousarray(is_word, np.uint8)
            desc = np.ascontiguousarray(desc, np.uint8).reshape(n, 32)
            weight = np.ascontiguousarray(weight, np.float64)
            _check("orb_vocabulary_create",
                   lib().orb_vocabulary_create(self.ctx.h, int(k), int(L), int(scoring), int(weighting), n,
                                               _ptr(parent), _ptr(is_word), _ptr(desc), _ptr(weight),
                                               ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().orb_vocabulary_destroy(self.h)
            self.h = None

    __del__ = close

    def info(self):
        v = [ctypes.c_int() for _ in range(6)]
        _check("orb_vocabulary_info", lib().orb_vocabulary_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("k", "L", "scoring", "weighting", "n_nodes", "n_words"), (x.value for x in v)))

    def transform(self, desc, levelsup=0):
        """Per descriptor: (word id, weight, node id at depth L - levelsup)."""
        desc = np.ascontiguousarray(desc, np.uint8).reshape(-1, 32)
        n = len(desc)
        wid = np.zeros(n, np.int32)
        w = np.zeros(n, np.float64)
        nid = np.zeros(n, np.int32)
        _check("orb_vocabulary_transform",
               lib().orb_vocabulary_transform(self.h, _ptr(desc), n, int(levelsup), _ptr(wid), _ptr(w), _ptr(nid)))
        return wid, w, nid

    def bow(self, images, levelsup=4):
        """transform(features, BowVector, FeatureVector, levelsup) for a list
        of N x 32 descriptor arrays: per image (words, values, fv_nodes,
        fv_features), in the reference's map order."""
        mats = [np.ascontiguousarray(d, np.uint8).reshape(-1, 32) for d in images]
        offs = np.zeros(len(mats) + 1, np.int32)
        offs[1:] = np.cumsum([len(d) for d in mats])
        rows = int(offs[-1])
        desc = np.concatenate(mats) if rows else np.zeros((0, 32), np.uint8)
        cap = max(rows, 1)
        bw, bv = np.zeros(cap, np.int32), np.zeros(cap, np.float64)
        fn, ff = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        nb, nf = np.zeros(len(mats), np.int32), np.zeros(len(mats), np.int32)
        _check("orb_vocabulary_bow",
               lib().orb_vocabulary_bow(self.h, len(mats), _ptr(desc), _ptr(offs), int(levelsup), _ptr(bw), _ptr(bv),
                                        _ptr(nb), _ptr(fn), _ptr(ff), _ptr(nf)))
        out = []
        for i in range(len(mats)):
            o = int(offs[i])
            out.append((bw[o:o + nb[i]].copy(), bv[o:o + nb[i]].copy(), fn[o:o + nf[i]].copy(),
                        ff[o:o + nf[i]].copy()))
        return out
