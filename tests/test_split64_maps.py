"""CPU restatement of the index maps of the FP64 colour-split V-cycle kernels
(split64.hip), checked against the oracle (no GPU).

The emulation follows the kernels thread by thread: `march_body` (two places
per thread, the x neighbours outside the pair taken from lane - 1 / lane + 1
of the same wave or loaded by the first and last lane; every value is loaded
here from its array, where the kernel keeps the other colour's z-neighbour
pairs in registers while it marches up the planes; MODE 0 plain, 1 zero
start, 2 prolongation folded in with a natural or colour-split coarse level,
3 zero start with b from the natural-order r, RZ writing z in natural order
with the (r, z) partials), `rr64_kernel` (red children only or all eight),
`split_h_kernel` and `reorder64_kernel`. Every global index is asserted in
range; a wrong place, wrap, colour or parent index shows up as a mismatch
with the oracle's GSRB sweep, residual, restriction and prolongation.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import bc_pair, random_K

DIRICHLET, NEUMANN, PERIODIC = 0, 1, 2


class Arr:
    """A flat array whose every read and write is bounds-checked."""

    def __init__(self, a):
        self.a = a

    def __getitem__(self, i):
        assert 0 <= i < self.a.size, i
        return self.a[i]

    def __setitem__(self, i, v):
        assert 0 <= i < self.a.size, i
        self.a[i] = v

    def ld2(self, i):
        assert i % 2 == 0 and 0 <= i and i + 1 < self.a.size, i
        return self.a[i], self.a[i + 1]

    def st2(self, i, u, v):
        assert i % 2 == 0 and 0 <= i and i + 1 < self.a.size, i
        self.a[i], self.a[i + 1] = u, v


def side(inside, in_off, t, wrap_off):
    nb = inside or t == PERIODIC
    face = nb or t == DIRICHLET
    off = in_off if inside else (wrap_off if t == PERIODIC else 0)
    return nb, face, off


def hface(sd, hc, hn):
    nb, face, _ = sd
    return 1.0 / (hc + (hn if nb else 0.0)) if face else 0.0


class Geo:
    def __init__(self, nx, ny, nz):
        assert nx % 4 == 0 and ny % 2 == 0 and nz % 2 == 0
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h, self.h2 = nx // 2, nx // 4
        self.half = self.h * ny * nz

    def at(self, i, j, k):
        return ((i + j + k) & 1) * self.half + (i >> 1) + self.h * (j + self.ny * k)


def to_split(G, nat):
    out = np.empty_like(nat)
    for k in range(G.nz):
        for j in range(G.ny):
            for i in range(G.nx):
                out[G.at(i, j, k)] = nat[i + G.nx * (j + G.ny * k)]
    return out


def to_nat(G, sp):
    out = np.empty_like(sp)
    for k in range(G.nz):
        for j in range(G.ny):
            for i in range(G.nx):
                out[i + G.nx * (j + G.ny * k)] = sp[G.at(i, j, k)]
    return out


def x_pair(O_, othc, base, r0, h, p, xlo, xhi, prv_lane, nxt_lane, per_lo, per_hi):
    a = O_.ld2(othc + base)
    prv = prv_lane if prv_lane is not None else (O_[othc + base - 1] if not xlo else None)
    nxt = nxt_lane if nxt_lane is not None else (O_[othc + base + 2] if not xhi else None)
    if p == 0:
        xp = [a[0], a[1]]
        xm = [(O_[othc + r0 + h - 1] if per_lo else 0.0) if xlo else prv, a[0]]
    else:
        xm = [a[0], a[1]]
        xp = [a[1], (O_[othc + r0] if per_hi else 0.0) if xhi else nxt]
    return a, xp, xm


def sweep(G, types, pin, bs, x, b, hs, colour, mode, tb=64, xc=None, cs=False, r64=None,
          z=None, rdot=None):
    """One half-sweep as march_body computes it (each plane of each column on
    its own); returns the per-thread-row partial sums (RZ)."""
    P = G.h2 * G.ny                 # threads of one plane
    tiles = -(-P // tb)             # workgroups per plane
    X, B, H = Arr(x), Arr(b), Arr(hs)
    per_lo, per_hi = types[0] == PERIODIC, types[1] == PERIODIC
    sy, sz = G.h, G.h * G.ny
    # a workgroup's lanes cover consecutive threads of one plane; lanes past
    # the plane mirror its last thread
    def thread_geo(k, tp):
        tq = min(tp, P - 1)
        q, j = tq % G.h2, tq // G.h2
        return tq, q, j + G.ny * k, j, k
    def pairs(k, tp):
        tq, q, row, j, _ = thread_geo(k, tp)
        base = row * G.h + 2 * q
        othc = (colour ^ 1) * G.half
        hp = H.ld2(othc + base)
        xp_ = X.ld2(othc + base) if mode in (0, 2) else (0.0, 0.0)
        R = None
        if mode == 2:
            R = rowpair(G, xc, cs, q, j >> 1, k >> 1)
        return hp, xp_, R
    cache = {(k, tp): pairs(k, tp) for k in range(G.nz) for tp in range(tiles * tb)}
    partials = np.zeros(tiles * G.nz)
    writes = []
    for k, tp in sorted(cache):
        tq, q, row, j, k = thread_geo(k, tp)
        t = (k, tp)
        valid = tp < P
        lane = tp % 64
        p = (j + k + colour) & 1
        r0 = row * G.h
        base = r0 + 2 * q
        own = colour * G.half + base
        othc = (colour ^ 1) * G.half
        oth = othc + base
        xlo, xhi = q == 0, q == G.h2 - 1
        Yp = side(j < G.ny - 1, sy, types[3], -(G.ny - 1) * sy)
        Ym = side(j > 0, -sy, types[2], (G.ny - 1) * sy)
        Zp = side(k < G.nz - 1, sz, types[5], -(G.nz - 1) * sz)
        Zm = side(k > 0, -sz, types[4], (G.nz - 1) * sz)
        Ex = [side(not (p == 1 and e == 1 and xhi), 0, types[1], 0) for e in range(2)]
        Wx = [side(not (p == 0 and e == 0 and xlo), 0, types[0], 0) for e in range(2)]
        hc = H.ld2(own)
        prv = cache[(k, tp - 1)][0][1] if lane > 0 else None
        nxt = cache[(k, tp + 1)][0][0] if lane < 63 and (k, tp + 1) in cache else None
        _, hxp, hxm = x_pair(H, othc, base, r0, G.h, p, xlo, xhi, prv, nxt, per_lo, per_hi)
        hy = [H.ld2(oth + Yp[2]), H.ld2(oth + Ym[2])]
        hz = [H.ld2(oth + Zp[2]), H.ld2(oth + Zm[2])]
        if mode in (1, 3):
            vxp = vxm = vyp = vym = vzp = vzm = [0.0, 0.0]
            xa = None
        else:
            prv = cache[(k, tp - 1)][1][1] if lane > 0 else None
            nxt = cache[(k, tp + 1)][1][0] if lane < 63 and (k, tp + 1) in cache else None
            xa, vxp, vxm = x_pair(X, othc, base, r0, G.h, p, xlo, xhi, prv, nxt, per_lo, per_hi)
            a, c = X.ld2(oth + Yp[2]), X.ld2(oth + Ym[2])
            u, w = X.ld2(oth + Zp[2]), X.ld2(oth + Zm[2])
            vyp = [v if Yp[0] else 0.0 for v in a]
            vym = [v if Ym[0] else 0.0 for v in c]
            vzp = [v if Zp[0] else 0.0 for v in u]
            vzm = [v if Zm[0] else 0.0 for v in w]
        vxp, vxm, vyp, vym, vzp, vzm = map(list, (vxp, vxm, vyp, vym, vzp, vzm))
        if mode == 3:
            R64 = Arr(r64)
            n01 = R64.ld2(row * G.nx + 4 * q)
            n23 = R64.ld2(row * G.nx + 4 * q + 2)
            nv = [n01[0], n01[1], n23[0], n23[1]]
            bb = [nv[p], nv[2 + p]]
            if valid:
                writes.append(("b", own, bb[0], bb[1]))
                writes.append(("b", oth, nv[1 - p], nv[3 - p]))
        else:
            bb = list(B.ld2(own))
        if mode == 2:
            NXc, NYc, NZc = G.nx >> 1, G.ny >> 1, G.nz >> 1
            J0, K0 = j >> 1, k >> 1
            R = cache[t][2]
            jodd, kodd = j & 1, k & 1
            Jp = (j + 1) >> 1 if j + 1 < G.ny else 0
            Jm = (j - 1) >> 1 if j > 0 else NYc - 1
            Kp = (k + 1) >> 1 if k + 1 < G.nz else 0
            Km = (k - 1) >> 1 if k > 0 else NZc - 1
            cy = rowpair(G, xc, cs, q, Jp if jodd else Jm, K0)
            cz = rowpair(G, xc, cs, q, J0, Kp if kodd else Km)
            for e in range(2):
                if Yp[0]: vyp[e] += cy[e] if jodd else R[e]
                if Ym[0]: vym[e] += R[e] if jodd else cy[e]
                if Zp[0]: vzp[e] += cz[e] if kodd else R[e]
                if Zm[0]: vzm[e] += R[e] if kodd else cz[e]
            rm = cache[(k, tp - 1)][2][1] if lane > 0 else (cval(G, xc, cs, 2 * q - 1, J0, K0) if not xlo else None)
            rpn = (cache[(k, tp + 1)][2][0] if lane < 63 and (k, tp + 1) in cache else
                   (cval(G, xc, cs, 2 * q + 2, J0, K0) if not xhi else None))
            if p == 0:
                vxp[0] += R[0]; vxp[1] += R[1]
                vxm[1] += R[0]
                if not xlo: vxm[0] += rm
                elif per_lo: vxm[0] += cval(G, xc, cs, NXc - 1, J0, K0)
            else:
                vxm[0] += R[0]; vxm[1] += R[1]
                vxp[0] += R[1]
                if not xhi: vxp[1] += rpn
                elif per_hi: vxp[1] += cval(G, xc, cs, 0, J0, K0)
        res = []
        for e in range(2):
            f = [hface(Ex[e], hc[e], hxp[e]), hface(Wx[e], hc[e], hxm[e]),
                 hface(Yp, hc[e], hy[0][e]), hface(Ym, hc[e], hy[1][e]),
                 hface(Zp, hc[e], hz[0][e]), hface(Zm, hc[e], hz[1][e])]
            v = [vxp[e], vxm[e], vyp[e], vym[e], vzp[e], vzm[e]]
            result = v[0] * f[0]
            for m in range(1, 6):
                result += v[m] * f[m]
            aC = f[0] + f[1] + f[2] + f[3] + f[4] + f[5]
            if pin and colour == 0 and row == 0 and q == 0 and e == 0:
                aC *= 2.0
            res.append((result - bb[e] * bs) / aC)
        if z is None:
            if valid:
                writes.append(("x", own, res[0], res[1]))
            continue
        zn = row * G.nx + 4 * q
        zz = [xa[0], res[0], xa[1], res[1]] if p else [res[0], xa[0], res[1], xa[1]]
        if valid:
            writes.append(("z", zn, zz[0], zz[1]))
            writes.append(("z", zn + 2, zz[2], zz[3]))
            ra, rb = Arr(rdot).ld2(zn), Arr(rdot).ld2(zn + 2)
            partials[k * tiles + tp // tb] += ra[0] * zz[0] + ra[1] * zz[1] + rb[0] * zz[2] + rb[1] * zz[3]
    # writes land after the reads of the whole half-sweep (no cell reads its own colour)
    for name, i, u, v in writes:
        {"x": X, "b": B, "z": Arr(z) if z is not None else None}[name].st2(i, u, v)
    return partials


def cval(G, xc, cs, I, J, K):
    NXc, NYc, NZc = G.nx >> 1, G.ny >> 1, G.nz >> 1
    C = Arr(xc)
    if cs:
        hcw = NXc >> 1
        return C[((I + J + K) & 1) * hcw * NYc * NZc + (I >> 1) + hcw * (J + NYc * K)]
    return C[I + NXc * (J + NYc * K)]


def rowpair(G, xc, cs, q, J, K):
    NXc, NYc, NZc = G.nx >> 1, G.ny >> 1, G.nz >> 1
    C = Arr(xc)
    if cs:
        hcw = NXc >> 1
        HC = hcw * NYc * NZc
        rb = q + hcw * (J + NYc * K)
        return [C[((J + K) & 1) * HC + rb], C[((J + K + 1) & 1) * HC + rb]]
    return list(C.ld2(2 * q + NXc * (J + NYc * K)))


def residual64(G, types, pin, X, B, H, i, j, k, rs):
    c = (i + j + k) & 1
    r0 = G.h * (j + G.ny * k)
    o = c * G.half + r0 + (i >> 1)
    Oo = (c ^ 1) * G.half
    sy, sz, m = G.h, G.h * G.ny, i >> 1
    sds = [side(i < G.nx - 1, r0 + ((i + 1) >> 1), types[1], r0),
           side(i > 0, r0 + ((i - 1) >> 1), types[0], r0 + G.h - 1),
           side(j < G.ny - 1, r0 + sy + m, types[3], r0 - (G.ny - 1) * sy + m),
           side(j > 0, r0 - sy + m, types[2], r0 + (G.ny - 1) * sy + m),
           side(k < G.nz - 1, r0 + sz + m, types[5], r0 - (G.nz - 1) * sz + m),
           side(k > 0, r0 - sz + m, types[4], r0 + (G.nz - 1) * sz + m)]
    xcv, hc = X[o], H[o]
    f, v = [], []
    for sd in sds:
        a = Oo + (sd[2] if sd[0] else r0)
        f.append(hface(sd, hc, H[a]))
        v.append(X[a] if sd[0] else 0.0)
    result = f[0] * (xcv - v[0])
    for m_ in range(1, 6):
        result += f[m_] * (xcv - v[m_])
    if pin and (i | j | k) == 0:
        result += sum(f) * xcv
    return B[o] + result * rs


def rr64(G, types, pin, rs, x, b, hs, red, cs):
    NXc, NYc, NZc = G.nx >> 1, G.ny >> 1, G.nz >> 1
    out = np.full(NXc * NYc * NZc, np.nan)
    X, B, H, Out = Arr(x), Arr(b), Arr(hs), Arr(out)
    for t in range(NXc * NYc * NZc):
        I, qq = t % NXc, t // NXc
        J, Kc = qq % NYc, qq // NYc
        i, j, k = 2 * I, 2 * J, 2 * Kc
        kids = ([(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)] if red else
                [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1),
                 (1, 1, 1)])
        s = 0.0
        for a, bb, c in kids:
            s += residual64(G, types, pin, X, B, H, i + a, j + bb, k + c, rs)
        oc = t
        if cs:
            hcw = NXc >> 1
            oc = ((I + J + Kc) & 1) * hcw * NYc * NZc + (I >> 1) + hcw * (J + NYc * Kc)
        Out[oc] = s / 8.0
    return out


CASES = [("darcy_x", (8, 6, 4)), ("periodic_yz", (12, 4, 6)), ("all_periodic", (8, 4, 4)),
         ("all_neumann", (16, 2, 4)), ("periodic_x_dirichlet_z", (8, 6, 6)),
         ("all_dirichlet", (4, 4, 2))]


def _setup(mix, shape, seed):
    og = O.grid(*shape, 0.5)
    types, _ = (lambda t: (t[0], t[1]))(__import__("helpers").BC_MIXES[mix])
    _, ob = bc_pair(mix)
    pin = mix in ("all_neumann", "all_periodic")
    n = O.ncells(og)
    rng = np.random.default_rng(seed)
    K = random_K(n, seed, 1.0)
    return og, ob, types, pin, n, rng, K


@pytest.mark.parametrize("tb", [64, 128])
@pytest.mark.parametrize("mix,shape", CASES)
def test_sweep_modes_0_1_3_and_rz_match_oracle(mix, shape, tb):
    og, ob, types, pin, n, rng, K = _setup(mix, shape, 3)
    G = Geo(*shape)
    hs = to_split(G, 0.5 * (1.0 / K))
    x0, rhs = rng.standard_normal(n), rng.standard_normal(n)
    bs = 0.25
    # MODE 0: one red-black iteration from x0
    x, b = to_split(G, x0), to_split(G, rhs)
    sweep(G, types, pin, bs, x, b, hs, 0, 0, tb)
    sweep(G, types, pin, bs, x, b, hs, 1, 0, tb)
    ref = O.gsrb_smooth_3d(og, ob, pin, x0, rhs, K, 1)
    assert np.abs(to_nat(G, x) - ref).max() < 1e-13 * np.abs(ref).max()
    # MODE 3 then RZ: zero start with b from natural r, z out with (r, z)
    xs = np.full(n, np.nan)
    bsplit = np.full(n, np.nan)
    sweep(G, types, pin, bs, xs, bsplit, hs, 0, 3, tb, r64=rhs)
    assert np.array_equal(bsplit, to_split(G, rhs))
    z = np.full(n, np.nan)
    part = sweep(G, types, pin, bs, xs, bsplit, hs, 1, 0, tb, z=z, rdot=rhs)
    ref0 = O.gsrb_smooth_3d(og, ob, pin, np.zeros(n), rhs, K, 1)
    assert np.abs(z - ref0).max() < 1e-13 * np.abs(ref0).max()
    assert abs(part.sum() - rhs @ ref0) < 1e-12 * np.abs(rhs) @ np.abs(ref0)
    # MODE 1 equals MODE 3 given the split b
    x1 = np.full(n, np.nan)
    sweep(G, types, pin, bs, x1, to_split(G, rhs), hs, 0, 1, tb)
    red = ((np.add.outer(np.add.outer(np.arange(shape[2]), np.arange(shape[1])),
                         np.arange(shape[0])).ravel()) & 1) == 0
    assert np.array_equal(to_nat(G, x1)[red], to_nat(G, xs)[red])


@pytest.mark.parametrize("cs", [False, True])
@pytest.mark.parametrize("mix,shape", [("darcy_x", (16, 4, 4)), ("all_periodic", (16, 8, 4)),
                                       ("periodic_yz", (8, 4, 8))])
def test_sweep_mode2_prolongation_matches_oracle(mix, shape, cs):
    og, ob, types, pin, n, rng, K = _setup(mix, shape, 5)
    G = Geo(*shape)
    coarse = (shape[0] // 2, shape[1] // 2, shape[2] // 2)
    if cs and coarse[0] % 4:
        pytest.skip("coarse level not split-capable")
    oc = O.grid(*coarse, 1.0)
    hs = to_split(G, 0.5 * (1.0 / K))
    x0, rhs = rng.standard_normal(n), rng.standard_normal(n)
    xcn = rng.standard_normal(O.ncells(oc))
    xc = to_split(Geo(*coarse), xcn) if cs else xcn
    x, b = to_split(G, x0), to_split(G, rhs)
    sweep(G, types, pin, 0.25, x, b, hs, 0, 2, 64, xc=xc, cs=cs)
    sweep(G, types, pin, 0.25, x, b, hs, 1, 0, 64)
    ref = O.gsrb_smooth_3d(og, ob, pin, O.prolong_3d_add(oc, xcn, x0), rhs, K, 1)
    assert np.abs(to_nat(G, x) - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("cs", [False, True])
@pytest.mark.parametrize("mix,shape", CASES)
def test_residual_restriction_matches_oracle(mix, shape, cs):
    og, ob, types, pin, n, rng, K = _setup(mix, shape, 7)
    G = Geo(*shape)
    coarse = (shape[0] // 2, shape[1] // 2, shape[2] // 2)
    if cs and (coarse[0] % 4 or coarse[1] % 2 or coarse[2] % 2):
        pytest.skip("coarse level not split-capable")
    oc = O.grid(*coarse, 1.0)
    hs = to_split(G, 0.5 * (1.0 / K))
    x0, rhs = rng.standard_normal(n), rng.standard_normal(n)
    rs = 1.0 / 0.25
    full = rr64(G, types, pin, rs, to_split(G, x0), to_split(G, rhs), hs, False, cs)
    ref = O.restrict_3d(oc, O.compute_residual_3d(og, ob, pin, x0, rhs, K))
    full_n = to_nat(Geo(*coarse), full) if cs else full
    assert np.abs(full_n - ref).max() < 1e-12 * np.abs(ref).max()
    # after a red-black sweep the black residuals vanish: red children suffice
    xs = O.gsrb_smooth_3d(og, ob, pin, x0, rhs, K, 1)
    redo = rr64(G, types, pin, rs, to_split(G, xs), to_split(G, rhs), hs, True, cs)
    ref = O.restrict_3d(oc, O.compute_residual_3d(og, ob, pin, xs, rhs, K))
    redo_n = to_nat(Geo(*coarse), redo) if cs else redo
    assert np.abs(redo_n - ref).max() < 1e-12 * np.abs(rhs).max() * 8


def test_planted_wrong_wrap_is_detected():
    """A periodic x wrap taken from the wrong place fails the oracle check."""
    og, ob, types, pin, n, rng, K = _setup("all_periodic", (8, 4, 4), 9)
    G = Geo(8, 4, 4)
    hs = to_split(G, 0.5 * (1.0 / K))
    x0, rhs = rng.standard_normal(n), rng.standard_normal(n)
    x, b = to_split(G, x0), to_split(G, rhs)
    global x_pair
    good = x_pair

    def bad(O_, othc, base, r0, h, p, xlo, xhi, prv, nxt, per_lo, per_hi):
        a, xp, xm = good(O_, othc, base, r0, h, p, xlo, xhi, prv, nxt, per_lo, per_hi)
        if p == 0 and xlo and per_lo:
            xm[0] = O_[othc + r0]  # place 0 instead of h - 1
        return a, xp, xm
    x_pair = bad
    try:
        sweep(G, types, pin, 0.25, x, b, hs, 0, 0)
        sweep(G, types, pin, 0.25, x, b, hs, 1, 0)
    finally:
        x_pair = good
    ref = O.gsrb_smooth_3d(og, ob, pin, x0, rhs, K, 1)
    assert np.abs(to_nat(G, x) - ref).max() > 1e-6
