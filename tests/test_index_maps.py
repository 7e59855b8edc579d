"""CPU restatements of the index maps of three HIP kernels, checked against
the oracle (no GPU): the plane-tile FP64 operator (`opz_lds_body`: LDS slot
L holds plane cell (t0 - nx + L) mod nx*ny, halo pairs dealt to threads,
idle pairs staging wrapped cells), the unit-stride r pass of the first
half-sweep (`split4_body` with RU: cell pairs 64 apart per lane, the own
colour's b handed over through LDS) and the division-free cell walk
(`CellWalk`). Every LDS slot and global index is asserted in range; a wrong
slot, wrap or colour shows up as a mismatch with the oracle's operator or
with the colour-split layout."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import BC_MIXES, bc_pair, grid_pair, random_K

KB = 256           # threads per workgroup (kBlock)
TILE = 4 * KB      # kOpTile: cells per plane tile
MAXNX = 512        # kOpLdsMaxNx
HALO = MAXNX // KB

DIRICHLET, NEUMANN, PERIODIC = 0, 1, 2


def _side(inside, t):
    nb = inside or t == PERIODIC
    return nb, nb or t == DIRICHLET


def _hface(s, hc, hn):
    nb, face = s
    return 1.0 / (hc + (hn if nb else 0.0)) if face else 0.0


def tile_operator(nx, ny, nz, types, pin, x, H, kz):
    """y = A x as opz_lds_body computes it, workgroup by workgroup."""
    P = nx * ny
    tiles = -(-P // TILE)
    chunks = -(-nz // kz)
    slots = TILE + 2 * nx
    y = np.full(nx * ny * nz, np.nan)
    for blk in range(tiles * chunks):
        chunk = blk // tiles
        t0 = (blk - chunk * tiles) * TILE
        k0, k1 = chunk * kz, min(nz, chunk * kz + kz)
        poff = lambda k: (k % nz) * P
        own = []
        for tid in range(KB):
            for m in range(2):
                c = t0 + 2 * (tid + m * KB)
                oc = c % P
                own.append((tid, m, c < P, oc, oc % nx, oc // nx, nx + 2 * (tid + m * KB)))
        halo = []
        for tid in range(KB):
            for m in range(HALO):
                q = tid + m * KB
                if q < nx:
                    L = 2 * q if q < nx // 2 else TILE + nx + 2 * (q - nx // 2)
                    halo.append((L, (t0 - nx + L) % P))
        lx = np.full(slots, np.nan)
        lh = np.full(slots, np.nan)

        def stage(k):
            po = poff(k)
            for (*_, oc, _i, _j, L) in own:
                assert 0 <= L and L + 1 < slots
                lx[L:L + 2] = x[po + oc:po + oc + 2]
                lh[L:L + 2] = H[po + oc:po + oc + 2]
            for L, c in halo:
                assert 0 <= L and L + 1 < slots
                lx[L:L + 2] = x[po + c:po + c + 2]
                lh[L:L + 2] = H[po + c:po + c + 2]

        stage(k0)
        for L in range(slots):  # the slot rule
            assert lx[L] == x[poff(k0) + (t0 - nx + L) % P]
        D0 = _side(k0 > 0, types[4])
        xm, fzl = {}, {}
        for (tid, m, act, oc, i, j, L) in own:
            po, pd = poff(k0), poff(k0 - 1)
            xd, hd, hc = x[pd + oc:pd + oc + 2], H[pd + oc:pd + oc + 2], H[po + oc:po + oc + 2]
            xm[tid, m] = xd.copy() if D0[0] else np.zeros(2)
            fzl[tid, m] = np.array([_hface(D0, hc[0], hd[0]), _hface(D0, hc[1], hd[1])])
        for k in range(k0, k1):
            if k > k0:
                stage(k)
            U = _side(k < nz - 1, types[5])
            for (tid, m, act, oc, i, j, L) in own:
                Wt, E = _side(i > 0, types[0]), _side(i + 2 < nx, types[1])
                N, S = _side(j < ny - 1, types[3]), _side(j > 0, types[2])
                po, pp = poff(k), poff(k + 1)
                xc, hc = x[po + oc:po + oc + 2], H[po + oc:po + oc + 2]
                Lw = L - 1 if i > 0 else L + nx - 1
                Le = L + 2 if i + 2 < nx else L + 2 - nx
                for s in (Lw, Le, L + nx, L + nx + 1, L - nx, L - nx + 1):
                    assert 0 <= s < slots
                xw, hw = (lx[Lw] if Wt[0] else 0.0), lh[Lw]
                xe, he = (lx[Le] if E[0] else 0.0), lh[Le]
                xn, hn, xs, hs = lx[L + nx:L + nx + 2], lh[L + nx:L + nx + 2], \
                    lx[L - nx:L - nx + 2], lh[L - nx:L - nx + 2]
                xp, hp = x[pp + oc:pp + oc + 2], H[pp + oc:pp + oc + 2]
                fw, fm, fe = _hface(Wt, hc[0], hw), 1.0 / (hc[0] + hc[1]), _hface(E, hc[1], he)
                fn = [_hface(N, hc[e], hn[e]) for e in (0, 1)]
                fs = [_hface(S, hc[e], hs[e]) for e in (0, 1)]
                fu = [_hface(U, hc[e], hp[e]) for e in (0, 1)]
                n = xn if N[0] else (0.0, 0.0)
                s = xs if S[0] else (0.0, 0.0)
                u = xp if U[0] else (0.0, 0.0)
                a, bz = xm[tid, m], fzl[tid, m]
                A0 = fw * (xc[0] - xw) + fm * (xc[0] - xc[1]) + fs[0] * (xc[0] - s[0]) + \
                    fn[0] * (xc[0] - n[0]) + bz[0] * (xc[0] - a[0]) + fu[0] * (xc[0] - u[0])
                A1 = fm * (xc[1] - xc[0]) + fe * (xc[1] - xe) + fs[1] * (xc[1] - s[1]) + \
                    fn[1] * (xc[1] - n[1]) + bz[1] * (xc[1] - a[1]) + fu[1] * (xc[1] - u[1])
                if pin and k == 0 and oc == 0 and act:
                    A0 += (fm + (fn[0] if ny > 1 else 0.0) + (fu[0] if nz > 1 else 0.0)) * xc[0]
                if act:
                    o = k * P + oc
                    assert np.isnan(y[o]), "cell written twice"
                    y[o], y[o + 1] = -A0, -A1
                xm[tid, m], fzl[tid, m] = xc.copy(), np.array(fu)
    assert not np.isnan(y).any(), "cell never written"
    return y


@pytest.mark.parametrize("mix", list(BC_MIXES))
@pytest.mark.parametrize("shape,kz", [((8, 6, 4), 1), ((10, 12, 3), 2), ((34, 34, 2), 1),
                                      ((64, 40, 2), 2), ((2, 2, 3), 1)])
def test_plane_tile_operator_index_map(shape, kz, mix):
    """Tiles smaller than a plane (several per plane, the last partial, rows
    straddling tiles) and planes smaller than a tile; z-chunks of 1 and 2."""
    g, og = grid_pair(*shape)
    _, obc = bc_pair(mix)
    types = BC_MIXES[mix][0]
    n = int(np.prod(shape))
    K = random_K(n, 5, 1.0)
    x = np.random.default_rng(6).standard_normal(n)
    pin = mix in ("all_neumann", "all_periodic")
    y = tile_operator(*shape, types, pin, x, 0.5 * (1.0 / K), kz)
    yo = O.varcoeff_apply(og, obc, pin, K, x)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()


def _split_index(nx, ny, nz, e):
    i, row = e % nx, e // nx
    j, k = row % ny, row // ny
    c = (i + j + k) & 1
    return c * (nx * ny * nz // 2) + row * (nx // 2) + i // 2


@pytest.mark.parametrize("shape", [(8, 2, 2), (16, 6, 4), (24, 10, 6), (256, 4, 2), (40, 2, 8)])
@pytest.mark.parametrize("tb", [64, 256, 512])
def test_unit_stride_r_pass_hands_each_thread_its_b(shape, tb):
    """The first half-sweep's r pass: lane l of a wave takes cell pairs
    (w0 + 128 u + 2 l, + 1), u < 4, writes both colours of b at the pair's
    colour-split place and the sweep colour's value to LDS slot
    64 u + l of the wave; the GS pass's thread t then finds the b of its four
    own places 4q .. 4q + 3 at slots 4t .. 4t + 3."""
    nx, ny, nz = shape
    n = nx * ny * nz
    half, h = n // 2, nx // 2
    r = np.random.default_rng(1).standard_normal(n)
    want = np.zeros(n)
    for e in range(n):
        want[_split_index(nx, ny, nz, e)] = r[e]
    total = n // 8
    for colour in (0, 1):
        got = np.full(n, np.nan)
        for blk in range(-(-total // tb)):
            bown = np.full(tb * 4, np.nan)
            for tid in range(tb):
                lane, w0 = tid & 63, 8 * (blk * tb + (tid & ~63))
                for u in range(4):
                    e = w0 + 128 * u + 2 * lane
                    if e < n:
                        row = e // nx
                        c0 = (row % ny + row // ny) & 1
                        got[c0 * half + e // 2] = r[e]
                        got[(c0 ^ 1) * half + e // 2] = r[e + 1]
                        slot = (tid & ~63) * 4 + 64 * u + lane
                        assert 0 <= slot < tb * 4
                        bown[slot] = r[e] if c0 == colour else r[e + 1]
            for tid in range(tb):
                t = blk * tb + tid
                if t < total:
                    q, row = t % (h // 4), t // (h // 4)
                    for e in range(4):
                        assert bown[4 * tid + e] == want[colour * half + row * h + 4 * q + e]
        assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [(8, 2, 2), (16, 6, 4), (24, 10, 6), (256, 4, 2), (8, 8, 8)])
def test_cell_walk_matches_division(shape):
    """CellWalk: (i, j, k) of e advanced by 128 cells at a time equals the
    divided-out indices of e + 128 u."""
    nx, ny, nz = shape
    n = nx * ny * nz
    for e0 in range(0, n, 2):
        i, row = e0 % nx, e0 // nx
        j, k = row % ny, row // ny
        for u in range(1, 4):
            i += 128
            while i >= nx:
                i -= nx
                j += 1
                if j == ny:
                    j, k = 0, k + 1
            e = e0 + 128 * u
            if e < n:
                assert (i, j, k) == (e % nx, (e // nx) % ny, e // (nx * ny))
