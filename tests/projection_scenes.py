"""Hand-built scenes and a second, independent statement of both per-frame
ORBmatcher::SearchByProjection overloads:

  SearchByProjection(Frame &F, const vector<MapPoint*> &vpMapPoints, th)
      ORBmatcher.cc:72-183, RadiusByViewingCos :186-202
  SearchByProjection(Frame &Current, const Frame &Last, th, bMono)
      ORBmatcher.cc:1710-1889, ComputeThreeMaxima :2035-2077

with the keypoint grid of Frame::AssignFeaturesToGrid / PosInGrid /
GetFeaturesInArea (Frame.cc:399-423, 590-700; this fork's circular window).
Every float expression is evaluated with numpy float32 scalars, one rounding
per operation, as the reference's unfused float code does.
"""
import math

import numpy as np

f32 = np.float32
TH_HIGH = 100
COLS, ROWS, HISTO = 64, 48, 30

KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"), ("response", "<f4"),
                     ("octave", "<i4"), ("class_id", "<i4")])
PROJ_DTYPE = np.dtype([("proj_x", "<f4"), ("proj_y", "<f4"), ("proj_xr", "<f4"), ("view_cos", "<f4"),
                       ("level", "<i4"), ("flags", "<i4")])
LAST_DTYPE = np.dtype([("xc", "<f4"), ("yc", "<f4"), ("zc", "<f4"), ("angle", "<f4"), ("octave", "<i4"),
                       ("flags", "<i4")])
IN_VIEW, BAD, HAS_OBS = 1, 2, 4
PRESENT = 1


def hamming(a, b):
    return int(np.unpackbits(np.bitwise_xor(a, b)).sum())


def _round_half_away(v):
    v = float(v)
    return int(math.copysign(math.floor(abs(v) + 0.5), v))


class FrameGrid:
    """Frame::AssignFeaturesToGrid: cells in index order (Frame.cc:399-423)."""

    def __init__(self, kp, bounds):
        self.kp = kp
        self.minX, self.maxX, self.minY, self.maxY = (f32(b) for b in bounds)
        self.invW = f32(f32(COLS) / f32(self.maxX - self.minX))
        self.invH = f32(f32(ROWS) / f32(self.maxY - self.minY))
        self.cells = {}
        for i, k in enumerate(kp):
            px = _round_half_away(f32(f32(f32(k["x"]) - self.minX) * self.invW))
            py = _round_half_away(f32(f32(f32(k["y"]) - self.minY) * self.invH))
            if 0 <= px < COLS and 0 <= py < ROWS:
                self.cells.setdefault((px, py), []).append(i)

    def area(self, x, y, r, min_level=-1, max_level=-1):
        """Frame::GetFeaturesInArea (Frame.cc:590-680)."""
        x, y, r = f32(x), f32(y), f32(r)
        cx0 = max(0, math.floor(f32(f32(f32(x - self.minX) - r) * self.invW)))
        if cx0 >= COLS:
            return []
        cx1 = min(COLS - 1, math.ceil(f32(f32(f32(x - self.minX) + r) * self.invW)))
        if cx1 < 0:
            return []
        cy0 = max(0, math.floor(f32(f32(f32(y - self.minY) - r) * self.invH)))
        if cy0 >= ROWS:
            return []
        cy1 = min(ROWS - 1, math.ceil(f32(f32(f32(y - self.minY) + r) * self.invH)))
        if cy1 < 0:
            return []
        check = min_level > 0 or max_level >= 0
        out = []
        for ix in range(cx0, cx1 + 1):
            for iy in range(cy0, cy1 + 1):
                for i in self.cells.get((ix, iy), []):
                    k = self.kp[i]
                    if check:
                        if k["octave"] < min_level:
                            continue
                        if max_level >= 0 and k["octave"] > max_level:
                            continue
                    dx, dy = f32(f32(k["x"]) - x), f32(f32(k["y"]) - y)
                    if f32(f32(dx * dx) + f32(dy * dy)) < f32(r * r):
                        out.append(i)
        return out


def search_local(kp, desc, uright, occupied, bounds, scale, pts, pdesc, th=1.0, nnratio=0.8):
    """SearchByProjection(F, vpMapPoints, th): (nmatches, pt_match, key_point)."""
    grid = FrameGrid(kp, bounds)
    owner = [-2 if o else -1 for o in occupied]          # -2: a map point with observations
    key_point = np.full(len(kp), -1, np.int32)
    pt_match = np.full(len(pts), -1, np.int32)
    th = f32(th)
    n = 0
    for p, P in enumerate(pts):
        if not P["flags"] & IN_VIEW or P["flags"] & BAD:
            continue
        lev = int(P["level"])
        r = f32(2.5) if float(f32(P["view_cos"])) > 0.998 else f32(4.0)
        if th != 1.0:
            r = f32(r * th)
        rad = f32(r * f32(scale[lev]))
        best, best_lev, best2, best_lev2, best_idx = 256, -1, 256, -1, -1
        for i in grid.area(P["proj_x"], P["proj_y"], rad, lev - 1, lev):
            o = owner[i]
            if o == -2 or (o >= 0 and pts[o]["flags"] & HAS_OBS):
                continue
            if uright[i] > 0 and f32(abs(f32(f32(P["proj_xr"]) - f32(uright[i])))) > rad:
                continue
            d = hamming(pdesc[p], desc[i])
            if d < best:
                best2, best_lev2 = best, best_lev
                best, best_lev, best_idx = d, int(kp[i]["octave"]), i
            elif d < best2:
                best2, best_lev2 = d, int(kp[i]["octave"])
        if best <= TH_HIGH:
            if best_lev == best_lev2 and f32(best) > f32(f32(nnratio) * f32(best2)):
                continue
            owner[best_idx] = p
            key_point[best_idx] = p
            pt_match[p] = best_idx
            n += 1
    return n, pt_match, key_point


def three_maxima(sizes):
    """ORBmatcher::ComputeThreeMaxima (:2035-2077) on bin sizes."""
    m1 = m2 = m3 = 0
    i1 = i2 = i3 = -1
    for i, s in enumerate(sizes):
        if s > m1:
            m3, m2, m1 = m2, m1, s
            i3, i2, i1 = i2, i1, i
        elif s > m2:
            m3, m2 = m2, s
            i3, i2 = i2, i
        elif s > m3:
            m3, i3 = s, i
    if m2 < f32(0.1) * f32(m1):
        i2 = i3 = -1
    elif m3 < f32(0.1) * f32(m1):
        i3 = -1
    return i1, i2, i3


def search_last(kp, desc, uright, occupied, bounds, scale, cam, mono, pts, pdesc, th, check_orientation=True):
    """SearchByProjection(Current, Last, th, bMono): (nmatches, pt_match, key_point)."""
    fx, fy, cx, cy, mbf, mb, tlc_z = (f32(c) for c in cam)
    forward = tlc_z > mb and not mono
    backward = -tlc_z > mb and not mono
    grid = FrameGrid(kp, bounds)
    owner = [-2 if o else -1 for o in occupied]
    key_point = np.full(len(kp), -1, np.int32)
    pt_match = np.full(len(pts), -1, np.int32)
    hist = [[] for _ in range(HISTO)]
    factor = f32(f32(HISTO) / f32(360.0))
    th = f32(th)
    n = 0
    for p, P in enumerate(pts):
        if not P["flags"] & PRESENT:
            continue
        invz = f32(1.0 / float(f32(P["zc"])))                # 1.0 / float in double, stored as float
        if invz < 0:
            continue
        u = f32(f32(f32(fx * f32(P["xc"])) * invz) + cx)
        v = f32(f32(f32(fy * f32(P["yc"])) * invz) + cy)
        if u < grid.minX or u > grid.maxX or v < grid.minY or v > grid.maxY:
            continue
        o = int(P["octave"])
        rad = f32(th * f32(scale[o]))
        if forward:
            cand = grid.area(u, v, rad, o)
        elif backward:
            cand = grid.area(u, v, rad, 0, o)
        else:
            cand = grid.area(u, v, rad, o - 1, o + 1)
        best, best_idx = 256, -1
        for i in cand:
            w = owner[i]
            if w == -2 or (w >= 0 and pts[w]["flags"] & HAS_OBS):
                continue
            if uright[i] > 0:
                ur = f32(u - f32(mbf * invz))
                if f32(abs(f32(ur - f32(uright[i])))) > rad:
                    continue
            d = hamming(pdesc[p], desc[i])
            if d < best:
                best, best_idx = d, i
        if best <= TH_HIGH:
            owner[best_idx] = p
            key_point[best_idx] = p
            pt_match[p] = best_idx
            n += 1
            if check_orientation:
                rot = f32(f32(P["angle"]) - f32(kp[best_idx]["angle"]))
                if rot < 0.0:
                    rot = f32(rot + f32(360.0))
                b = _round_half_away(f32(rot * factor))
                if b == HISTO:
                    b = 0
                hist[b].append(best_idx)
    if check_orientation:
        keep = three_maxima([len(h) for h in hist])
        for b in range(HISTO):
            if b not in keep:
                for i in hist[b]:
                    key_point[i] = -2
                    n -= 1
    return n, pt_match, key_point


# ------------------------------------------------------------------ scenes
SCALE = np.array([f32(1.2) ** k for k in range(8)], np.float32)
BOUNDS = (0.0, 640.0, 0.0, 480.0)


class Scene:
    """A frame (keypoints, descriptors, mvuRight, occupancy) and the points
    projected into it."""

    def __init__(self, seed, last=False):
        self.rng = np.random.default_rng(seed)
        self.kp, self.desc, self.uright, self.occ = [], [], [], []
        self.pts, self.pdesc = [], []
        self.last = last
        # fx fy cx cy mbf mb tlc_z; with fx = 512 and z a power of 2 the
        # pinhole arithmetic is exact for integer pixel positions
        self.cam = (512.0, 512.0, 320.0, 240.0, 40.0, 0.08, 0.0)
        self.mono = False

    def rand_desc(self):
        return self.rng.integers(0, 256, 32).astype(np.uint8)

    def near(self, d, bits):
        e = d.copy()
        for b in self.rng.choice(256, bits, replace=False):
            e[b // 8] ^= np.uint8(1 << (b % 8))
        return e

    def key(self, x, y, octave=0, desc=None, uright=-1.0, occupied=False, angle=0.0):
        self.kp.append((x, y, 31.0 * SCALE[octave], angle, 20.0, octave, -1))
        self.desc.append(self.rand_desc() if desc is None else desc)
        self.uright.append(uright)
        self.occ.append(1 if occupied else 0)
        return len(self.kp) - 1

    def point(self, desc, **f):
        if self.last:
            rec = (f.get("xc", 0.0), f.get("yc", 0.0), f.get("zc", 1.0), f.get("angle", 0.0), f.get("octave", 0),
                   f.get("flags", PRESENT | HAS_OBS))
        else:
            rec = (f["x"], f["y"], f.get("xr", -1.0), f.get("view_cos", 0.9), f.get("level", 0),
                   f.get("flags", IN_VIEW | HAS_OBS))
        self.pts.append(rec)
        self.pdesc.append(desc)
        return len(self.pts) - 1

    def camera_point(self, desc, u, v, z, **f):
        """A last-frame point that projects to (u, v); exactly so for integer
        u, v and z a power of 2."""
        fx, fy, cx, cy = self.cam[:4]
        return self.point(desc, xc=(u - cx) * z / fx, yc=(v - cy) * z / fy, zc=z, **f)

    def arrays(self):
        kp = np.array(self.kp, KP_DTYPE)
        desc = np.array(self.desc, np.uint8).reshape(-1, 32)
        pts = np.array(self.pts, LAST_DTYPE if self.last else PROJ_DTYPE)
        pdesc = np.array(self.pdesc, np.uint8).reshape(-1, 32)
        return (kp, desc, np.array(self.uright, np.float32), np.array(self.occ, np.uint8), np.array(BOUNDS, np.float32),
                SCALE, pts, pdesc)


def scene_radius():
    """RadiusByViewingCos: 2.5 when viewCos > 0.998 (a double compare, so the
    float 0.998f counts as above), else 4.0; a key 3 px away is found only
    with 4.0, one 2.4 px away with both."""
    s = Scene(1)
    for i, vc in enumerate((0.999, 0.99, np.float32(0.998), 0.9)):
        x, y = 60.0 + 80 * i, 100.0
        d = s.rand_desc()
        s.key(x + 3.0, y, 0, s.near(d, 10))
        s.point(d, x=x, y=y, view_cos=vc)
        x2 = x + 40
        d2 = s.rand_desc()
        s.key(x2 + 2.4, y, 0, s.near(d2, 10))
        s.point(d2, x=x2, y=y, view_cos=vc)
    return s


def scene_levels():
    """Window levels [level - 1, level]; ratio only between same-level best
    and second; TH_HIGH accepted at 100, rejected at 101."""
    s = Scene(2)
    d = s.rand_desc()
    for lev, bits in ((0, 1), (1, 20), (2, 22), (3, 2)):     # only levels 1 and 2 are searched
        s.key(100.0 + lev * 0.5, 100.0, lev, s.near(d, bits))
    s.point(d, x=101.0, y=100.0, level=2)                  # best 20 @1, second 22 @2: different levels -> kept
    d = s.rand_desc()
    s.key(200.0, 100.0, 1, s.near(d, 20))
    s.key(200.5, 100.0, 1, s.near(d, 22))
    s.point(d, x=200.0, y=100.0, level=1)                  # same level, 20 > 0.8 * 22: rejected
    d = s.rand_desc()
    s.key(300.0, 100.0, 1, s.near(d, 10))
    s.key(300.5, 100.0, 1, s.near(d, 22))
    s.point(d, x=300.0, y=100.0, level=1)                  # 10 <= 0.8 * 22: kept
    for k, bits in enumerate((100, 101)):
        d = s.rand_desc()
        s.key(400.0 + 60 * k, 100.0, 0, s.near(d, bits))
        s.point(d, x=400.0 + 60 * k, y=100.0, level=0)
    return s


def scene_window():
    """A key inside the square but outside the circle of radius r, one on
    the circle (distance^2 == r^2, excluded) and one just inside."""
    s = Scene(3)
    d = s.rand_desc()
    s.key(102.0, 102.0, 0, s.near(d, 5))                   # (2, 2): inside square, 8 > 4^2? no -> inside circle
    s.key(103.0, 103.0, 0, s.near(d, 3))                   # (3, 3): 18 > 16 -> outside circle, inside square
    s.point(d, x=100.0, y=100.0, view_cos=0.9)             # r = 4
    d = s.rand_desc()
    s.key(204.0, 200.0, 0, s.near(d, 3))                   # distance exactly 4: not < 16
    s.key(196.5, 200.0, 0, s.near(d, 30))
    s.point(d, x=200.0, y=200.0, view_cos=0.9)
    return s


def scene_claims():
    """Keypoints held by map points with observations are skipped; a claim in
    this call blocks later points only if its point has observations."""
    s = Scene(4)
    d = s.rand_desc()
    s.key(100.0, 100.0, 0, s.near(d, 2), occupied=True)
    s.key(101.0, 100.0, 0, s.near(d, 30))
    s.point(d, x=100.0, y=100.0)                           # takes the free key at 30
    d = s.rand_desc()
    k_a = s.key(200.0, 100.0, 0, s.near(d, 5))
    s.key(201.0, 100.0, 0, s.near(d, 40))
    s.point(d, x=200.0, y=100.0)                           # claims k_a (has observations)
    s.point(s.near(d, 1), x=200.0, y=100.0)                # k_a is blocked: falls back to the 40-bit key
    d = s.rand_desc()
    s.key(300.0, 100.0, 0, s.near(d, 5))
    s.point(d, x=300.0, y=100.0, flags=IN_VIEW)            # no observations: its claim does not block
    s.point(s.near(d, 2), x=300.0, y=100.0)                # takes the same key; key_point is overwritten
    d = s.rand_desc()
    s.key(400.0, 100.0, 0, s.near(d, 5))
    s.point(d, x=400.0, y=100.0, flags=0)                  # not in view
    s.point(d, x=400.0, y=100.0, flags=IN_VIEW | BAD | HAS_OBS)
    assert k_a >= 0
    return s


def scene_stereo():
    """er = |mTrackProjXR - mvuRight| against r * scale: above is rejected,
    equal is kept; mvuRight <= 0 is not checked."""
    s = Scene(5)
    for i, (xr, ur) in enumerate(((50.0, 54.5), (50.0, 54.0), (50.0, -1.0), (50.0, 0.0))):
        d = s.rand_desc()
        s.key(100.0 + 100 * i, 200.0, 0, s.near(d, 3), uright=ur)
        s.point(d, x=100.0 + 100 * i, y=200.0, xr=xr, view_cos=0.9)    # r = 4
    return s


def scene_th():
    """th != 1 scales the radius (run with th = 3): a key 10 px away is found."""
    s = Scene(6)
    d = s.rand_desc()
    s.key(110.0, 100.0, 0, s.near(d, 4))
    s.point(d, x=100.0, y=100.0, view_cos=0.9)             # r = 4 * 3 = 12
    d = s.rand_desc()
    s.key(313.0, 100.0, 0, s.near(d, 4))
    s.point(d, x=300.0, y=100.0, view_cos=0.9)             # 13 > 12: not found
    return s


def scene_last_levels(direction):
    """Last-frame points projected exactly; the level window follows the
    motion: forward -> levels >= o (no check at all when o = 0), backward ->
    [0, o], otherwise [o - 1, o + 1]."""
    s = Scene(7, last=True)
    s.cam = (512.0, 512.0, 320.0, 240.0, 40.0, 0.08, {1: 0.5, -1: -0.5, 0: 0.0}[direction])
    for j, o in enumerate((0, 2, 4)):
        d = s.rand_desc()
        for lev in range(8):
            s.key(100.0 + 150 * j + lev * 0.25, 100.0 + 40 * lev, lev, s.near(d, 40 - lev * 3 if lev != o else 30))
        for lev in range(8):
            s.camera_point(d, 100.0 + 150 * j + lev * 0.25, 100.0 + 40 * lev, 2.0, octave=o, angle=0.0)
    return s


def scene_last_projection():
    """Behind the camera (1/z < 0), outside the bounds, exact projection, the
    stereo check against u - mbf / z."""
    s = Scene(8, last=True)
    d = s.rand_desc()
    s.key(100.0, 100.0, 0, s.near(d, 5))
    s.camera_point(d, 100.0, 100.0, -2.0)                  # behind
    s.camera_point(s.near(d, 1), 100.0, 100.0, 2.0)        # found
    s.camera_point(d, -1.0, 100.0, 2.0)                    # u < mnMinX
    for i, off in enumerate((0.0, 15.0, 15.5)):            # th 15 at level 0 -> radius 15
        d = s.rand_desc()
        u = 200.0 + 100 * i
        ur = u - 40.0 / 2.0 + off                          # ur = u - mbf * invz = u - 20
        s.key(u, 300.0, 0, s.near(d, 5), uright=ur)
        s.camera_point(d, u, 300.0, 2.0)
    return s


def scene_last_rotation():
    """The rotation histogram keeps the three fullest bins (the third only at
    >= 10 % of the first); a rotation of exactly 6 deg falls in bin 1 (half
    away from zero); claims by points without observations do not block."""
    s = Scene(9, last=True)
    k = 0
    for count, rot in ((12, 30.0), (6, 90.0), (2, 150.0), (3, 6.0), (1, 200.0)):
        for _ in range(count):
            d = s.rand_desc()
            u, v = 40.0 + 50 * (k % 12), 40.0 + 60 * (k // 12)
            s.key(u, v, 0, s.near(d, 4), angle=100.0)
            s.camera_point(d, u, v, 4.0, angle=100.0 + rot)
            k += 1
    d = s.rand_desc()
    s.key(600.0, 400.0, 0, s.near(d, 4), angle=10.0)
    s.camera_point(d, 600.0, 400.0, 4.0, angle=40.0, flags=PRESENT)   # no observations
    s.camera_point(s.near(d, 1), 600.0, 400.0, 4.0, angle=10.0)       # claims it too (rot 0: bin 0)
    return s


LOCAL_SCENES = [("radius", scene_radius, 1.0), ("levels", scene_levels, 1.0), ("window", scene_window, 1.0),
                ("claims", scene_claims, 1.0), ("stereo", scene_stereo, 1.0), ("th", scene_th, 3.0)]
LAST_SCENES = [("forward", lambda: scene_last_levels(1), 15.0, False),
               ("backward", lambda: scene_last_levels(-1), 15.0, False),
               ("still", lambda: scene_last_levels(0), 15.0, False),
               ("mono", lambda: scene_last_levels(1), 15.0, True),
               ("projection", scene_last_projection, 15.0, False),
               ("rotation", scene_last_rotation, 15.0, False)]


def random_frame(seed, n_kp=1500, n_pts=1200, last=False, w=640, h=480):
    """A dense random scene: keypoints over 8 levels, points projected near
    them (most within a few px of a keypoint whose descriptor is a noisy copy
    of the point's), some occupied keys and stereo keys."""
    rng = np.random.default_rng(seed)
    s = Scene(seed, last=last)
    kx = rng.uniform(0, w, n_kp).astype(np.float32)
    ky = rng.uniform(0, h, n_kp).astype(np.float32)
    lev = rng.choice(8, n_kp, p=[0.3, 0.2, 0.15, 0.1, 0.09, 0.07, 0.05, 0.04])
    kd = rng.integers(0, 256, (n_kp, 32)).astype(np.uint8)
    ang = rng.uniform(0, 360, n_kp).astype(np.float32)
    ur = np.where(rng.random(n_kp) < 0.5, kx - rng.uniform(0, 40, n_kp).astype(np.float32), -1.0)
    for i in range(n_kp):
        s.key(float(kx[i]), float(ky[i]), int(lev[i]), kd[i], uright=float(ur[i]), occupied=rng.random() < 0.05,
              angle=float(ang[i]))
    for p in range(n_pts):
        t = int(rng.integers(0, n_kp))
        d = s.near(kd[t], int(rng.integers(0, 60)))
        x = float(kx[t] + rng.normal(0, 3))
        y = float(ky[t] + rng.normal(0, 3))
        if last:
            s.cam = (512.0, 512.0, 320.0, 240.0, 40.0, 0.08, float(seed % 3 - 1) * 0.5)
            z = float(rng.uniform(1.0, 20.0))
            if rng.random() < 0.02:
                z = -z
            fl = PRESENT | (HAS_OBS if rng.random() < 0.9 else 0) if rng.random() < 0.95 else 0
            s.camera_point(d, x, y, z, octave=int(lev[t]), angle=float((ang[t] + rng.normal(0, 8)) % 360), flags=fl)
        else:
            fl = (IN_VIEW if rng.random() < 0.95 else 0) | (BAD if rng.random() < 0.02 else 0) | \
                 (HAS_OBS if rng.random() < 0.95 else 0)
            s.point(d, x=x, y=y, xr=float(x - rng.uniform(0, 40)), view_cos=float(rng.uniform(0.99, 1.0)),
                    level=int(np.clip(lev[t] + rng.integers(-1, 2), 0, 7)), flags=fl)
    return s
