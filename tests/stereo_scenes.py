"""Hand-built stereo scenes and a second, independent statement of
Frame::ComputeStereoMatches (Frame.cc:831-1082), for the known-answer tests.

The statement below is written from the reference text in plain Python with
numpy float32 scalars for every float expression, so that each float
operation rounds exactly once, as the reference's unfused float code does.

Each scene places keypoints so that one rule of the reference decides their
outcome:
  * the row band of the right keypoints (:867-882) and the truncated left row
    (:907);
  * the octave window (:929), the closed uR window [uL - bf/b, uL] (:936);
  * the strict '<' best from TH_HIGH and the '< 75' acceptance (:918-947, :950);
  * this fork's iniu = scaleduR0 - L - w and endu bounds (:992-997);
  * rejection of a best shift at +-L (:1023);
  * the parabola (:1031-1038) - |deltaR| > 1 (:1040-1044) cannot happen: with
    integer SADs and the first-minimum rule, d1 > d2 <= d3 gives |deltaR| <= 1/2;
  * disparity exactly 0 -> 0.01 and uL - 0.01 in double (:1051-1054);
  * the median cut at 1.5f*1.4f*median with a SAD exactly at the threshold
    (:1069-1081), and even-sized sets where the median index matters.
"""
import math

import numpy as np

f32 = np.float32
TH_HIGH, TH_LOW = 100, 50


def hamming(a, b):
    return int(np.unpackbits(np.bitwise_xor(a, b)).sum())


def _round_half_away(v):
    v = float(v)
    return math.copysign(math.floor(abs(v) + 0.5), v)


def compute_stereo_matches(levL, levR, scale, inv_scale, kl, dl, kr, dr, bf, b):
    """levL/levR: lists of u8 level images; kl/kr: keypoint records (x, y,
    octave); dl/dr: N x 32 u8.  Returns (uRight, depth) as float32 arrays."""
    n = len(kl)
    uRight = np.full(n, -1, np.float32)
    depth = np.full(n, -1, np.float32)
    th_orb = (TH_HIGH + TH_LOW) // 2
    n_rows = levL[0].shape[0]

    rows = [[] for _ in range(n_rows)]
    for iR in range(len(kr)):
        y = f32(kr[iR]["y"])
        r = f32(f32(2.0) * f32(scale[int(kr[iR]["octave"])]))
        hi = math.ceil(float(f32(y + r)))
        lo = math.floor(float(f32(y - r)))
        for yi in range(lo, hi + 1):
            if 0 <= yi < n_rows:
                rows[yi].append(iR)

    min_z = f32(b)
    min_d = f32(0)
    max_d = f32(f32(bf) / min_z)
    kept = []
    for iL in range(n):
        lev = int(kl[iL]["octave"])
        vL, uL = f32(kl[iL]["y"]), f32(kl[iL]["x"])
        row = int(vL)                       # float -> size_t truncates
        if row >= n_rows:
            continue
        cands = rows[row]
        if not cands:
            continue
        min_u, max_u = f32(uL - max_d), f32(uL - min_d)
        if max_u < 0:
            continue
        best, best_r = TH_HIGH, 0
        for iR in cands:
            o = int(kr[iR]["octave"])
            if o < lev - 1 or o > lev + 1:
                continue
            uR = f32(kr[iR]["x"])
            if min_u <= uR <= max_u:
                d = hamming(dl[iL], dr[iR])
                if d < best:
                    best, best_r = d, iR
        if not best < th_orb:
            continue
        inv = f32(inv_scale[lev])
        su = int(_round_half_away(f32(uL * inv)))
        sv = int(_round_half_away(f32(vL * inv)))
        su0 = f32(_round_half_away(f32(f32(kr[best_r]["x"]) * inv)))
        w = L = 5
        iniu = f32(f32(su0 - f32(L)) - f32(w))
        endu = f32(f32(f32(su0 + f32(L)) + f32(w)) + f32(1))
        if iniu < 0 or endu >= levR[lev].shape[1]:
            continue
        IL = levL[lev][sv - w:sv + w + 1, su - w:su + w + 1].astype(np.float32)
        IL = IL - IL[w, w]
        best_sad, best_inc = 2 ** 31 - 1, 0
        dists = [f32(0)] * (2 * L + 1)
        for inc in range(-L, L + 1):
            c = int(su0) + inc
            IR = levR[lev][sv - w:sv + w + 1, c - w:c + w + 1].astype(np.float32)
            IR = IR - IR[w, w]
            dist = f32(np.abs(IL - IR).astype(np.float64).sum())
            if dist < best_sad:
                best_sad, best_inc = int(dist), inc
            dists[L + inc] = dist
        if best_inc in (-L, L):
            continue
        d1, d2, d3 = dists[L + best_inc - 1], dists[L + best_inc], dists[L + best_inc + 1]
        delta = f32(f32(d1 - d3) / f32(f32(2.0) * f32(f32(d1 + d3) - f32(f32(2.0) * d2))))
        if delta < -1 or delta > 1:
            continue
        best_uR = f32(f32(scale[lev]) * f32(f32(su0 + f32(best_inc)) + delta))
        disp = f32(uL - best_uR)
        if disp >= min_d and disp < max_d:
            if disp <= 0:
                disp = f32(0.01)
                best_uR = f32(float(uL) - 0.01)
            depth[iL] = f32(f32(bf) / disp)
            uRight[iL] = best_uR
            kept.append((best_sad, iL))
    if kept:
        kept.sort()
        median = f32(kept[len(kept) // 2][0])
        th = f32(f32(f32(1.5) * f32(1.4)) * median)
        for sad, iL in reversed(kept):
            if sad < th:
                break
            uRight[iL] = -1
            depth[iL] = -1
    return uRight, depth


# ------------------------------------------------------------------ scenes
W, H = 480, 120
SHIFT = 12          # default disparity: a left point at u appears at u - 12


class Scene:
    """Left/right images plus hand-placed keypoints and descriptors.

    The left image is uniform noise in [30, 190) (so a SAD bump of up to 60
    never saturates); the right image shows it SHIFT px to the left, except
    where a region overrides that."""

    def __init__(self, seed, bf=386.1, b=0.5372):
        self.rng = np.random.default_rng(seed)
        self.bf, self.b = bf, b
        self.left = self.rng.integers(30, 190, (H, W)).astype(np.uint8)
        self.right = self.left.copy()
        self.right[:, :W - SHIFT] = self.left[:, SHIFT:]
        self.kl, self.kr, self.dl, self.dr = [], [], [], []

    def shift_region(self, x0, x1, d):
        """Right columns [x0, x1) show the left image d px further right:
        R(y, x) = L(y, x + d), so a left point at u appears at u - d."""
        self.right[:, x0:x1] = self.left[:, x0 + d:x1 + d]

    def mirror_region(self, xl, half, d):
        """Make the left image mirror-symmetric about column xl over +-half,
        and show it in the right image d px to the left (about xl - d)."""
        for k in range(1, half + 1):
            self.left[:, xl - k] = self.left[:, xl + k]
        self.right[:, xl - d - half:xl - d + half + 1] = self.left[:, xl - half:xl + half + 1]

    def desc(self):
        return self.rng.integers(0, 256, 32).astype(np.uint8)

    def flipped(self, d, distance):
        e = d.copy()
        for bit in self.rng.choice(256, distance, replace=False):
            e[bit // 8] ^= np.uint8(1 << (bit % 8))
        return e

    def add_left(self, x, y, octave=0, desc=None):
        self.kl.append((x, y, octave))
        self.dl.append(self.desc() if desc is None else desc)
        return len(self.kl) - 1

    def add_right(self, x, y, octave=0, desc=None):
        self.kr.append((x, y, octave))
        self.dr.append(self.desc() if desc is None else desc)
        return len(self.kr) - 1

    def pair(self, xl, y, xr=None, octave=0, octave_r=None, distance=0, yr=None, sad=10):
        """A left keypoint and a right keypoint (default: at xl - SHIFT) whose
        descriptors differ in exactly `distance` bits.  An octave-0 pair's
        matching window gets SAD `sad` (10 by default, so that a scene's
        median SAD is 10 and the outlier cut keeps SADs up to 20)."""
        d = self.desc()
        xr = xl - SHIFT if xr is None else xr
        il = self.add_left(xl, y, octave, d)
        ir = self.add_right(xr, y if yr is None else yr, octave if octave_r is None else octave_r,
                            self.flipped(d, distance))
        if sad and octave == 0:
            self.bump_sad(int(round(xr)), int(round(y)), sad)
        return il, ir

    def bump_sad(self, xr, y, k):
        """Raise the right image at (xr, y + 1) by k.  Every SAD window of a
        right keypoint at (xr, y) within +-5 px holds that pixel, off its
        centre row: the matching window's SAD grows by exactly k, and a
        mirror-symmetric scene stays symmetric (SAD(-1) = SAD(+1))."""
        self.right[y + 1, xr] = np.uint8(int(self.right[y + 1, xr]) + k)

    def keypoints(self, dtype):
        def pack(kps):
            a = np.zeros(len(kps), dtype)
            for i, (x, y, o) in enumerate(kps):
                a[i]["x"], a[i]["y"], a[i]["octave"] = x, y, o
                a[i]["size"], a[i]["class_id"] = 31.0, -1
            return a
        dl = np.array(self.dl, np.uint8).reshape(-1, 32)
        dr = np.array(self.dr, np.uint8).reshape(-1, 32)
        return pack(self.kl), dl, pack(self.kr), dr


def scene_bounds():
    """iniu = scaleduR0 - 10 and endu = scaleduR0 + 11 at and past the edges."""
    s = Scene(11)
    s.pair(22.0, 20.0, 10.0)            # iniu = 0: kept
    s.pair(21.0, 40.0, 9.0)             # iniu = -1: skipped
    s.shift_region(W - 40, W - 3, 3)
    s.pair(W - 9.0, 60.0, W - 12.0)     # endu = W - 1: kept
    s.pair(W - 8.0, 80.0, W - 11.0)     # endu = W: skipped
    s.pair(200.0, 100.0)                # well inside: kept
    return s


def scene_shift_limits():
    """The true correspondence at shifts -5, -4, 4, 5 of the right keypoint."""
    s = Scene(12)
    s.shift_region(100, 140, 7)         # left u -> right u - 7
    s.shift_region(200, 240, 9)
    s.shift_region(300, 340, 6)
    s.shift_region(400, 440, 8)
    s.pair(127.0, 30.0, 120.0 + 5)      # best at inc -5: rejected
    s.pair(229.0, 50.0, 220.0 + 4)      # best at -4: kept
    s.pair(326.0, 70.0, 320.0 - 4)      # best at +4: kept
    s.pair(428.0, 90.0, 420.0 - 5)      # best at +5: rejected
    s.pair(60.0, 60.0)                  # an ordinary match beside them
    return s


def scene_zero_disparity():
    """Mirror-symmetric texture about the keypoint: SAD(-1) = SAD(+1), so
    deltaR = 0 and the refined uR lands on an integer."""
    s = Scene(13, bf=10.0, b=0.5)       # maxD = 20 exactly
    s.mirror_region(150, 16, 0)
    s.pair(150.0, 40.0, 150.0)          # disparity exactly 0: the 0.01 step
    s.mirror_region(300, 16, 0)
    s.pair(300.0, 80.0, 300.0)
    s.mirror_region(400, 16, 20)
    s.pair(400.0, 60.0, 380.0)          # disparity exactly maxD: rejected
    s.pair(60.0, 60.0)                  # an ordinary positive disparity beside them
    s.pair(230.0, 100.0)
    return s


def scene_median(sads, seed):
    """One clean match per entry of `sads`, each with its SAD set to that value."""
    s = Scene(seed)
    for i, k in enumerate(sads):
        x, y = 40 + 40 * (i % 10), 20 + 30 * (i // 10)
        s.pair(float(x), float(y), sad=k)
    return s


def scene_hamming():
    """Descriptor rules: 74 kept, 75 rejected, 100 never accepted; a tie
    between two right candidates goes to the lower index; octave +-1; the uR
    window [uL - bf/b, uL] is closed at both ends."""
    s = Scene(14, bf=10.0, b=0.5)       # maxD = 20 exactly
    s.pair(40.0, 20.0, distance=74)
    s.pair(80.0, 20.0, distance=75)
    s.pair(120.0, 20.0, distance=100)
    # tie at 10 bits: a wrong candidate 7 px left (lower index) and the true one
    d = s.desc()
    s.add_left(200.0, 40.0, 0, d)
    s.add_right(181.0, 40.0, 0, s.flipped(d, 10))
    s.add_right(188.0, 40.0, 0, s.flipped(d, 10))
    # a nearer descriptor at a farther octave is ignored; octave +1 counts
    s.pair(260.0, 60.0, octave=0, octave_r=2, distance=0)
    s.pair(300.0, 60.0, octave=0, octave_r=1, distance=5)
    # uR > uL is outside the window
    s.pair(340.0, 80.0, 340.5, distance=0)
    # a closer descriptor exactly at uL - maxD is a candidate (and wins), one
    # 0.5 px farther is not
    d = s.desc()
    s.add_left(400.0, 100.0, 0, d)
    s.add_right(380.0, 100.0, 0, s.flipped(d, 3))
    s.add_right(388.0, 100.0, 0, s.flipped(d, 10))
    d = s.desc()
    s.add_left(440.0, 60.0, 0, d)
    s.add_right(419.5, 60.0, 0, s.flipped(d, 3))
    s.add_right(428.0, 60.0, 0, s.flipped(d, 10))
    return s


def scene_rows():
    """Row band floor(y - 2s) .. ceil(y + 2s) of the right keypoint, and the
    truncated left row."""
    s = Scene(15)
    s.pair(60.0, 52.9, yr=50.0)         # row 52 is in [48, 52]
    s.pair(120.0, 53.0, yr=50.0)        # row 53 is not
    s.pair(180.0, 52.6, yr=50.0)        # truncates to 52 (rounding would give 53)
    s.pair(240.0, 47.2, yr=50.0)        # row 47 is not
    s.pair(300.0, 48.0, yr=50.0)        # row 48 is
    s.pair(360.0, 70.0, octave=1, octave_r=1, yr=67.7)  # band +-2.4 at octave 1
    s.pair(420.0, 90.0)
    return s


def scene_octave1():
    """Octave-1 keypoints: coordinates scaled to level 1 and rounded."""
    s = Scene(16)
    for i in range(8):
        x, y = 60.0 + 45 * i + 0.37 * i, 30.0 + 7 * i
        s.pair(x, y, octave=1)
    for i in range(9):                  # octave-0 pairs at SAD 10 set the median
        s.pair(40.0 + 48 * i, 104.0)
    return s


# (name, builder) pairs; the median scenes carry SAD sets whose cut depends on
# an exact threshold tie (median 10 -> th = 21.0) or on the median's index.
SCENES = [
    ("bounds", scene_bounds),
    ("shift_limits", scene_shift_limits),
    ("zero_disparity", scene_zero_disparity),
    ("median_tie", lambda: scene_median([3, 10, 10, 10, 10, 20, 21, 40], 21)),
    ("median_index", lambda: scene_median([3, 4, 10, 12, 25, 26, 30, 60], 22)),
    ("median_zero", lambda: scene_median([0, 0, 0, 5, 9], 23)),
    ("hamming", scene_hamming),
    ("rows", scene_rows),
    ("octave1", scene_octave1),
]
