"""Seeded synthetic inputs for the ORB front end.

Every workload used by the tests and the benchmark is generated here from an
integer seed with numpy's PCG64 generator, so the GPU path and the CPU oracle
see byte-identical images.

* ``shapes_frame``   -- random axis-aligned rectangles and ellipses on a
  random background, plus uniform noise, Gaussian-blurred (sigma 1) and
  clipped to uint8.
* ``warp_affine``    -- rotation about the image centre plus a translation,
  bilinear, border clamped (the second frame of the small plain workload and
  the camera-stream sequences).
* ``stereo_pair``    -- a left frame and the right view obtained through a
  piecewise-planar disparity map (Voronoi regions of random seeds, one plane
  per region, disparity clipped to [0, max_disp]).
* ``descriptor_sets``-- uniform random 256-bit queries and a map set with
  planted near-duplicates (Binomial(256, p) flipped bits).
* ``stream_frames``  -- per-camera sequences: each frame is the previous one
  rotated U[0,2] deg, shifted U[0,8] px, plus fresh uniform noise.
"""
import numpy as np


def _gauss_kernel(sigma):
    r = int(np.ceil(3 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum(), r


def _blur(img, sigma):
    k, r = _gauss_kernel(sigma)
    p = np.pad(img, r, mode="reflect")
    h, w = img.shape
    tmp = np.zeros((h + 2 * r, w), np.float64)
    for i, c in enumerate(k):
        tmp += c * p[:, i:i + w]
    out = np.zeros((h, w), np.float64)
    for i, c in enumerate(k):
        out += c * tmp[i:i + h, :]
    return out


def shapes_frame(width, height, seed, n_shapes=200, noise=20.0, sigma=1.0):
    """One uint8 frame of random rectangles/ellipses plus blurred noise."""
    rng = np.random.default_rng(seed)
    img = np.full((height, width), rng.uniform(0, 255), np.float64)
    for _ in range(n_shapes):
        ellipse = bool(rng.integers(0, 2))
        cx, cy = rng.uniform(0, width), rng.uniform(0, height)
        rx = rng.uniform(4, max(8.0, width / 8))
        ry = rng.uniform(4, max(8.0, height / 8))
        val = rng.uniform(0, 255)
        x0, x1 = max(0, int(cx - rx)), min(width, int(cx + rx) + 1)
        y0, y1 = max(0, int(cy - ry)), min(height, int(cy + ry) + 1)
        if x0 >= x1 or y0 >= y1:
            continue
        if ellipse:
            yy, xx = np.mgrid[y0:y1, x0:x1]
            m = ((xx - cx) / rx) ** 2 + ((yy - cy) / ry) ** 2 <= 1.0
            img[y0:y1, x0:x1][m] = val
        else:
            img[y0:y1, x0:x1] = val
    img += rng.uniform(-noise, noise, size=(height, width))
    img = _blur(img, sigma)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def _bilinear(img, xs, ys):
    h, w = img.shape
    xs = np.clip(xs, 0, w - 1)
    ys = np.clip(ys, 0, h - 1)
    x0 = np.floor(xs).astype(np.int64)
    y0 = np.floor(ys).astype(np.int64)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    fx = xs - x0
    fy = ys - y0
    f = img.astype(np.float64)
    top = f[y0, x0] * (1 - fx) + f[y0, x1] * fx
    bot = f[y1, x0] * (1 - fx) + f[y1, x1] * fx
    return top * (1 - fy) + bot * fy


def warp_affine(img, angle_deg, tx, ty):
    """Rotate by angle_deg about the centre and translate by (tx, ty)."""
    h, w = img.shape
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    t = np.deg2rad(angle_deg)
    c, s = np.cos(t), np.sin(t)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    u, v = xx - cx - tx, yy - cy - ty
    xs = c * u + s * v + cx
    ys = -s * u + c * v + cy
    out = _bilinear(img, xs, ys)
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def disparity_map(width, height, seed, n_planes=12, max_disp=64.0):
    """Piecewise-planar disparity: one plane per Voronoi region of n seeds."""
    rng = np.random.default_rng(seed)
    sx = rng.uniform(0, width, n_planes)
    sy = rng.uniform(0, height, n_planes)
    d0 = rng.uniform(0, max_disp, n_planes)
    gx = rng.uniform(-0.05, 0.05, n_planes)
    gy = rng.uniform(-0.05, 0.05, n_planes)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    dist = (xx[None] - sx[:, None, None]) ** 2 + (yy[None] - sy[:, None, None]) ** 2
    region = np.argmin(dist, axis=0)
    disp = d0[region] + gx[region] * (xx - sx[region]) + gy[region] * (yy - sy[region])
    return np.clip(disp, 0.0, max_disp)


def stereo_pair(width, height, seed, n_planes=12, max_disp=64.0):
    """(left, right, disparity): right(x, y) = left(x + d(x, y), y)."""
    left = shapes_frame(width, height, seed)
    disp = disparity_map(width, height, seed + 7919, n_planes, max_disp)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    right = np.clip(np.rint(_bilinear(left, xx + disp, yy)), 0, 255).astype(np.uint8)
    return left, right, disp


def frame_pair(width, height, seed, angle_deg=3.0, shift=(6.0, 4.0)):
    """A frame and its copy rotated by angle_deg and shifted by `shift`."""
    a = shapes_frame(width, height, seed)
    return a, warp_affine(a, angle_deg, shift[0], shift[1])


def descriptor_sets(n_query, n_map, seed, planted_frac=0.1, flip_p=0.08):
    """Random 256-bit queries and a map set with planted near-duplicates.

    Returns (queries[n_query, 32], map[n_map, 32], planted_query[n_map]) where
    planted_query[j] is the query a planted map row was derived from, or -1.
    """
    rng = np.random.default_rng(seed)
    q = rng.integers(0, 256, size=(n_query, 32), dtype=np.uint8)
    m = rng.integers(0, 256, size=(n_map, 32), dtype=np.uint8)
    n_plant = int(round(planted_frac * n_map))
    rows = rng.choice(n_map, size=n_plant, replace=False)
    src = rng.integers(0, n_query, size=n_plant)
    flips = rng.random((n_plant, 256)) < flip_p
    mask = np.packbits(flips, axis=1, bitorder="little")
    m[rows] = q[src] ^ mask
    planted = np.full(n_map, -1, np.int64)
    planted[rows] = src
    return q, m, planted


def stream_frames(width, height, seed, n_frames, max_rot=2.0, max_shift=8.0, noise=10.0):
    """A camera sequence: each frame is the previous one moved plus noise."""
    rng = np.random.default_rng(seed + 104729)
    cur = shapes_frame(width, height, seed)
    frames = [cur]
    for _ in range(n_frames - 1):
        moved = warp_affine(cur, rng.uniform(0, max_rot), rng.uniform(0, max_shift),
                            rng.uniform(0, max_shift)).astype(np.float64)
        moved += rng.uniform(-noise, noise, size=moved.shape)
        cur = np.clip(np.rint(moved), 0, 255).astype(np.uint8)
        frames.append(cur)
    return frames


def _sparse_bits(rng, shape, ands):
    """Random bytes whose bits are set with probability 2**-ands."""
    out = rng.integers(0, 256, size=shape, dtype=np.uint8)
    for _ in range(ands - 1):
        out &= rng.integers(0, 256, size=shape, dtype=np.uint8)
    return out


def vocabulary(k=10, L=6, seed=0, flip_ands=3, stop_frac=0.02, scoring=0, weighting=0):
    """A DBoW2-shaped vocabulary tree in file order, breadth first: node 0 is
    the root, the k**l nodes of level l follow level l-1, node j's children
    are consecutive.  Level-1 descriptors are uniform; a child is its parent
    with each bit flipped with probability 2**-flip_ands.  Leaves are the
    words, with idf-like weights in [0.5, 8) and a stop_frac share of zero
    weights (stopped words).  Returns a dict of the constructor arguments of
    Vocabulary / OracleVocabulary."""
    rng = np.random.default_rng(seed + 7919)
    parents = [np.zeros(1, np.int32)]
    descs = [np.zeros((1, 32), np.uint8)]
    prev = np.zeros(1, np.int64)
    first = 1
    for level in range(1, L + 1):
        n = len(prev) * k
        par = np.repeat(prev, k).astype(np.int32)
        if level == 1:
            d = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
        else:
            d = np.repeat(descs[-1], k, axis=0) ^ _sparse_bits(rng, (n, 32), flip_ands)
        parents.append(par)
        descs.append(d)
        prev = np.arange(first, first + n, dtype=np.int64)
        first += n
    parent = np.concatenate(parents)
    desc = np.concatenate(descs)
    nn = len(parent)
    is_word = np.zeros(nn, np.uint8)
    is_word[nn - k ** L:] = 1
    weight = np.zeros(nn, np.float64)
    nw = k ** L
    w = rng.uniform(0.5, 8.0, size=nw)
    w[rng.random(nw) < stop_frac] = 0.0
    weight[nn - nw:] = w
    return dict(k=k, L=L, scoring=scoring, weighting=weighting, parent=parent, is_word=is_word, desc=desc,
                weight=weight)


def vocabulary_queries(voc, n, seed, flip_ands=4):
    """n descriptors near random words of a vocabulary() tree: a word's
    descriptor with each bit flipped with probability 2**-flip_ands."""
    rng = np.random.default_rng(seed + 104723)
    words = np.flatnonzero(voc["is_word"])
    pick = voc["desc"][rng.choice(words, size=n)]
    return pick ^ _sparse_bits(rng, pick.shape, flip_ands)


PROJ_POINT_DTYPE = np.dtype([("proj_x", "<f4"), ("proj_y", "<f4"), ("proj_xr", "<f4"), ("view_cos", "<f4"),
                             ("level", "<i4"), ("flags", "<i4")])
LAST_POINT_DTYPE = np.dtype([("xc", "<f4"), ("yc", "<f4"), ("zc", "<f4"), ("angle", "<f4"), ("octave", "<i4"),
                             ("flags", "<i4")])
KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"), ("response", "<f4"),
                           ("octave", "<i4"), ("class_id", "<i4")])


def projection_frame(seed, n_kp=2000, n_pts=1800, last=False, width=640, height=480):
    """A tracking-like SearchByProjection input: keypoints over 8 levels
    (half with mvuRight, 5 % already holding a map point) and points
    projected a few px from a keypoint whose descriptor is a noisy copy of
    theirs.  Returns (kp, desc, uright, occupied, bounds, points, point_desc,
    camera) with points in the record of the overload (last=True: camera
    coordinates for fx = fy = 512, cx = 320, cy = 240, mbf = 40)."""
    rng = np.random.default_rng(seed)
    scale = np.float32(1.2) ** np.arange(8, dtype=np.float32)
    kp = np.zeros(n_kp, KEYPOINT_DTYPE)
    kp["x"] = rng.uniform(0, width, n_kp)
    kp["y"] = rng.uniform(0, height, n_kp)
    kp["octave"] = rng.choice(8, n_kp, p=[0.3, 0.2, 0.15, 0.1, 0.09, 0.07, 0.05, 0.04])
    kp["size"] = 31.0 * scale[kp["octave"]]
    kp["angle"] = rng.uniform(0, 360, n_kp)
    kp["class_id"] = -1
    desc = rng.integers(0, 256, (n_kp, 32), dtype=np.uint8)
    uright = np.where(rng.random(n_kp) < 0.5, kp["x"] - rng.uniform(0, 40, n_kp), -1.0).astype(np.float32)
    occupied = (rng.random(n_kp) < 0.05).astype(np.uint8)
    t = rng.integers(0, n_kp, n_pts)
    flips = rng.integers(0, 60, n_pts)
    pdesc = desc[t].copy()
    for p in range(n_pts):
        bits = rng.choice(256, flips[p], replace=False)
        np.bitwise_xor.at(pdesc[p], bits // 8, (1 << (bits % 8)).astype(np.uint8))
    x = kp["x"][t] + rng.normal(0, 3, n_pts)
    y = kp["y"][t] + rng.normal(0, 3, n_pts)
    cam = (512.0, 512.0, 320.0, 240.0, 40.0, 0.08, 0.5 * (seed % 3 - 1))
    if last:
        pts = np.zeros(n_pts, LAST_POINT_DTYPE)
        z = rng.uniform(1.0, 20.0, n_pts)
        pts["xc"] = (x - cam[2]) * z / cam[0]
        pts["yc"] = (y - cam[3]) * z / cam[1]
        pts["zc"] = z
        pts["angle"] = (kp["angle"][t] + rng.normal(0, 8, n_pts)) % 360
        pts["octave"] = kp["octave"][t]
        pts["flags"] = np.where(rng.random(n_pts) < 0.95, 1 | 4, 0)
    else:
        pts = np.zeros(n_pts, PROJ_POINT_DTYPE)
        pts["proj_x"], pts["proj_y"] = x, y
        pts["proj_xr"] = x - rng.uniform(0, 40, n_pts)
        pts["view_cos"] = rng.uniform(0.99, 1.0, n_pts)
        pts["level"] = np.clip(kp["octave"][t] + rng.integers(-1, 2, n_pts), 0, 7)
        pts["flags"] = np.where(rng.random(n_pts) < 0.95, 1 | 4, 0)
    return kp, desc, uright, occupied, (0.0, float(width), 0.0, float(height)), pts, pdesc, cam
