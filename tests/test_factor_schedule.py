"""CPU restatement of the batched coarsest factor (`band_factor_kernel` in
stencil.hip; no GPU), thread by thread and barrier interval by barrier
interval: the cell-parallel band assembly (`band_assemble`), the LDS-window
schedule (`band_factor_window`: wave 0 updates and scales column j + 1,
waves 1..15 update columns j + 2 .. j + w from a fixed pair table, the last
two waves store the column entering the window into the slot column j - 1
left) and the global-memory variant (`band_factor_global`).

Checked on the coarsest shapes of every boundary mix and on random banded
matrices up to the window's limit (w + 1 = 90):
  * the assembled band equals the oracle's smoother/residual operator (the
    operator the coarsest level solves, pin included) column by column;
  * within one barrier interval no LDS (or, for the global variant, C)
    address that an operation writes is read or written by any other
    operation, so the interval's order of execution cannot matter;
  * the pair table gives every update (dk, t) to exactly one thread;
  * every entry of the row-major factor is written exactly once;
  * the window schedule's factor equals the global variant's bit for bit
    (same operations on every entry, in the same order), L L^T equals the
    oracle's operator to 1e-13, and both variants flag the same first pivot
    that is not positive.
The fused multiply-add of the kernels is modelled as c - a * l in both
variants, so the bit-for-bit comparison checks the schedule, not the FMA."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import BC_MIXES, bc_pair, grid_pair, random_K

NT = 1024          # kFactorThreads
WIN = 90           # kFactorWin: largest w + 1 of the LDS window
PAIRS = 5          # kFactorPairs
UPD = NT - 64      # threads of waves 1..15
DIRICHLET, PERIODIC = 0, 2


def neighbour(nx, ny, nz, types, i, j, k, d):
    """Linear index of the neighbour across face d, -1 at a closed boundary."""
    ijk, dims = [i, j, k], (nx, ny, nz)
    ax, up = d // 2, d & 1
    at = ijk[ax] == (dims[ax] - 1 if up else 0)
    if at and types[d] != PERIODIC:
        return -1
    ijk[ax] = (0 if up else dims[ax] - 1) if at else ijk[ax] + (1 if up else -1)
    return ijk[0] + nx * (ijk[1] + ny * ijk[2])


def bandwidth(nx, ny, nz, types):
    w = 0
    for c in range(nx * ny * nz):
        i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
        for d in range(6):
            q = neighbour(nx, ny, nz, types, i, j, k, d)
            if q >= 0:
                w = max(w, abs(q - c))
    return w


def assemble(nx, ny, nz, types, pin, K, dx2, w):
    """band_assemble: column c of the lower band, entry t = A[c + t][c]."""
    n, W1 = nx * ny * nz, w + 1
    C = np.zeros((n, W1))
    for c in range(n):
        i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
        nb, coef, asum = [], [], 0.0
        for d in range(6):
            q = neighbour(nx, ny, nz, types, i, j, k, d)
            if q < 0:
                cf = 2.0 * K[c] if types[d] == DIRICHLET else 0.0
            else:
                cf = 2.0 / (1.0 / K[c] + 1.0 / K[q])
            nb.append(q)
            coef.append(cf)
            asum += cf
        if pin and c == 0:
            asum *= 2.0
        e = [-cf / dx2 for cf in coef]
        diag = asum / dx2
        for d in range(6):
            if nb[d] == c:
                diag += e[d]
        for d in range(6):
            if nb[d] <= c or nb[d] in nb[:d]:
                continue
            acc = e[d]
            for d2 in range(d + 1, 6):
                if nb[d2] == nb[d]:
                    acc += e[d2]
            assert nb[d] - c <= w
            C[c, nb[d] - c] = acc
        C[c, 0] = diag
    return C


def dense_operator(shape, mix, K):
    """A from the oracle's residual: at a pinned cell 0 its diagonal doubles
    the sum over all six faces (the operator of the whole V-cycle)."""
    g, og = grid_pair(*shape)
    _, obc = bc_pair(mix)
    n = int(np.prod(shape))
    pin = mix in ("all_neumann", "all_periodic")
    A, zero = np.zeros((n, n)), np.zeros(n)
    for q in range(n):
        e = np.zeros(n)
        e[q] = 1.0
        A[:, q] = O.compute_residual_3d(og, obc, pin, e, zero, K)  # r = 0 - (-A) e_q
    return A, pin


def band_to_dense(C):
    n, W1 = C.shape
    A = np.zeros((n, n))
    for c in range(n):
        for t in range(min(W1, n - c)):
            A[c + t, c] = A[c, c + t] = C[c, t]
    return A


class Interval:
    """The accesses of one barrier interval, each tagged with its operation."""

    def __init__(self):
        self.r_op, self.r_addr, self.w_op, self.w_addr = [], [], [], []

    def read(self, op, addr):
        self.r_op.append(np.broadcast_to(np.asarray(op), np.shape(addr)).ravel())
        self.r_addr.append(np.asarray(addr).ravel())

    def write(self, op, addr):
        self.w_op.append(np.asarray(op).ravel())
        self.w_addr.append(np.asarray(addr).ravel())

    def check(self, size):
        cat = lambda a: np.concatenate(a) if a else np.zeros(0, dtype=np.int64)
        wa, wo = cat(self.w_addr), cat(self.w_op)
        ra, ro = cat(self.r_addr), cat(self.r_op)
        assert np.all((wa >= 0) & (wa < size)) and np.all((ra >= 0) & (ra < size))
        assert np.unique(wa).size == wa.size, "an address written twice in one interval"
        owner = np.full(size, -1)
        owner[wa] = wo
        hit = owner[ra]
        assert np.all((hit < 0) | (hit == ro)), "an address read and written by two operations"


def pair_table(w):
    """The (dk, t) pairs of the update threads: pair r of thread tid is the
    flat entry f = tid - 64 + r * 960 of the list dk = 2..w, t = 0..w - dk."""
    W1 = w + 1
    pdk = np.zeros((NT, PAIRS), dtype=np.int64)
    pt = np.zeros((NT, PAIRS), dtype=np.int64)
    lens = np.array([W1 - dk for dk in range(2, w + 1)], dtype=np.int64)
    e0 = np.concatenate(([0], np.cumsum(lens)))
    tid = np.arange(64, NT)
    for r in range(PAIRS):
        f = tid - 64 + r * UPD
        slot = np.searchsorted(e0, f, side="right") - 1
        ok = f < e0[-1]
        pdk[tid[ok], r] = slot[ok] + 2
        pt[tid[ok], r] = f[ok] - e0[slot[ok]]
    return pdk, pt


def factor_window(C, w):
    """band_factor_window: returns (Lb, write counts of Lb, flag)."""
    n, W1, WS = C.shape[0], w + 1, w + 2
    assert W1 <= WIN and w >= 1
    Cf = C.ravel()
    win = np.zeros(WS * W1)
    for c in range(min(W1, n)):
        win[c * W1:(c + 1) * W1] = C[c]
    Lb = np.full(n * W1, np.nan)
    cnt = np.zeros(n * W1, dtype=np.int64)
    for e in range(W1 * W1):
        i, d = divmod(e, W1)
        if d > i and i < n:
            Lb[e] = 0.0
            cnt[e] += 1

    pdk, pt = pair_table(w)
    used = pdk[64:] > 0
    flat = (pdk[64:][used] - 2) * W1 + pt[64:][used]
    assert np.unique(flat).size == flat.size == w * (w - 1) // 2, "pair table"
    assert np.all(pt[64:][used] + pdk[64:][used] <= w)

    bad = 0
    m0 = min(w, n - 1)
    t = np.arange(m0 + 1)
    L0 = np.sqrt(win[0])
    if not win[0] > 0.0:
        bad = 1
    v = np.where(t == 0, L0, win[t] / L0)
    win[t[1:]] = v[1:]
    Lb[t * W1 + t] = v
    cnt[t * W1 + t] += 1

    pf = np.arange(W1)                     # entry of the loader thread NT - 128 + pf
    assert W1 <= 128
    pre = Cf[(w + 1) * W1 + pf].copy() if w + 1 < n else None
    sj = 0
    lane = np.arange(64)
    for j in range(n):
        m = min(w, n - 1 - j)
        cj = sj * W1
        s1 = 0 if sj + 1 == WS else sj + 1
        iv = Interval()
        new_addr, new_val = [], []
        if m >= 1:                          # wave 0: column j + 1, op id = lane
            m1 = min(w, n - 2 - j)
            assert m1 <= 127
            c1 = s1 * W1
            l1 = win[cj + 1]
            piv = None
            vals = {}
            for tt in (lane, lane + 64):
                live = tt <= m1
                x = tt[live]
                iv.read(lane[live], c1 + x)
                iv.read(lane[live], np.full(x.size, cj + 1))
                vv = win[c1 + x].copy()
                upd = x <= m - 1
                iv.read(lane[live][upd], cj + 1 + x[upd])
                vv[upd] = vv[upd] - win[cj + 1 + x[upd]] * l1
                vals[int(tt[0])] = (x, vv, lane[live])
            x0, v0, _ = vals[0]
            piv = v0[0]
            L = np.sqrt(piv)
            if not piv > 0.0 and bad == 0:
                bad = j + 2
            for x, vv, ops in vals.values():
                out = np.where(x == 0, L, vv / L)
                keep = x > 0
                iv.write(ops[keep], c1 + x[keep])
                new_addr.append(c1 + x[keep])
                new_val.append(out[keep])
                rows = (j + 1 + x) * W1 + x
                assert np.all(np.isnan(Lb[rows]))
                Lb[rows] = out
                cnt[rows] += 1
        # loaders (waves 14 and 15), op id = 64 + loader entry
        if j + 1 + w < n:
            dst = (WS - 1 if sj == 0 else sj - 1) * W1 + pf
            iv.write(64 + pf, dst)
            new_addr.append(dst)
            new_val.append(pre)
        if j + 2 + w < n:
            pre = Cf[(j + 2 + w) * W1 + pf].copy()
        # waves 1..15: op id = 64 + W1 + pair index
        dk, tt = pdk[64:].ravel(), pt[64:].ravel()
        ok = (dk > 0) & (dk + tt <= m)
        dk, tt = dk[ok], tt[ok]
        ops = 64 + W1 + np.flatnonzero(ok)
        s = (sj + dk) % WS
        at = s * W1 + tt
        iv.read(ops, cj + dk + tt)
        iv.read(ops, cj + dk)
        iv.read(ops, at)
        iv.write(ops, at)
        new_addr.append(at)
        new_val.append(win[at] - win[cj + dk + tt] * win[cj + dk])
        iv.check(WS * W1)
        for a, vv in zip(new_addr, new_val):
            win[a] = vv
        sj = s1
    return Lb, cnt, bad


def factor_global(C, w):
    """band_factor_global: the factor in place in C, two barriers per step."""
    n, W1 = C.shape[0], w + 1
    Cf = C.ravel().copy()
    Lb = np.full(n * W1, np.nan)
    for e in range(W1 * W1):
        i, d = divmod(e, W1)
        if d > i and i < n:
            Lb[e] = 0.0
    bad = 0
    for j in range(n):
        m = min(w, n - 1 - j)
        cj = j * W1
        djj = Cf[cj]
        Ljj = np.sqrt(djj)
        if not djj > 0.0 and bad == 0:
            bad = j + 1
        t = np.arange(m + 1)
        iv = Interval()
        iv.read(t, np.full(t.size, cj))
        iv.read(t, cj + t)
        iv.write(t[1:], cj + t[1:])
        iv.check(n * W1)
        v = np.where(t == 0, Ljj, Cf[cj + t] / Ljj)
        Cf[cj + t[1:]] = v[1:]
        Lb[(j + t) * W1 + t] = v
        iv = Interval()
        dk = np.repeat(np.arange(1, m + 1), [m - d + 1 for d in range(1, m + 1)])
        tt = np.concatenate([np.arange(m - d + 1) for d in range(1, m + 1)]) if m else dk
        ops = np.arange(dk.size)
        ck = (j + dk) * W1 + tt
        iv.read(ops, cj + dk + tt)
        iv.read(ops, cj + dk)
        iv.read(ops, ck)
        iv.write(ops, ck)
        iv.check(n * W1)
        Cf[ck] = Cf[ck] - Cf[cj + dk + tt] * Cf[cj + dk]
    return Lb, bad


def lower_from_band(Lb, n, W1):
    L = np.zeros((n, n))
    for i in range(n):
        for d in range(min(W1, i + 1)):
            L[i, i - d] = Lb[i * W1 + d]
    return L


def check_factor(C, w, A):
    n, W1 = C.shape[0], w + 1
    Lw, cnt, badw = factor_window(C, w)
    Lg, badg = factor_global(C, w)
    assert np.all(cnt == 1), "a factor entry written twice or never"
    assert badw == badg == 0
    assert np.array_equal(Lw, Lg)
    L = lower_from_band(Lw, n, W1)
    assert np.abs(L @ L.T - A).max() <= 1e-13 * np.abs(A).max()


SHAPES = [
    ((4, 4, 4), "darcy_x"),                 # w = 16
    ((2, 3, 4), "all_neumann"),             # pinned, w = 6
    ((2, 2, 2), "all_periodic"),            # pinned, coinciding x neighbours
    ((5, 4, 1), "all_periodic"),            # z faces wrap onto the cell itself
    ((6, 5, 2), "periodic_yz"),             # w = 30
    ((4, 4, 5), "periodic_x_dirichlet_z"),  # w = 16
    ((9, 9, 3), "all_dirichlet"),           # w = 81
    ((89, 2, 1), "darcy_x"),                # w = 89: the widest window
]


@pytest.mark.parametrize("shape,mix", SHAPES)
def test_window_schedule_factors_the_coarsest_operator(shape, mix):
    types = BC_MIXES[mix][0]
    n = int(np.prod(shape))
    K = random_K(n, 11, 1.0)
    A, pin = dense_operator(shape, mix, K)
    w = bandwidth(*shape, types)
    assert 1 <= w and w + 1 <= WIN
    C = assemble(*shape, types, pin, K, 1.0, w)
    assert np.abs(band_to_dense(C) - A).max() <= 1e-14 * np.abs(A).max()
    check_factor(C, w, A)


@pytest.mark.parametrize("seed", range(6))
def test_window_schedule_on_random_bands(seed):
    """Diagonally dominant bands with random sparsity, n from 2 to 300 and w
    up to the window's limit, including n <= w + 1 (the window never fills)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 301)) if seed else 40
    w = int(rng.integers(1, min(n - 1, WIN - 1) + 1)) if seed else 39
    C = np.where(rng.random((n, w + 1)) < 0.4, -rng.random((n, w + 1)), 0.0)
    C[:, 1:][np.arange(n)[:, None] + np.arange(1, w + 1) >= n] = 0.0
    A = band_to_dense(np.c_[np.zeros(n), C[:, 1:]])
    C[:, 0] = np.abs(A).sum(axis=0) + 1.0
    A = band_to_dense(C)
    check_factor(C, w, A)


def test_both_variants_flag_the_first_pivot_that_is_not_positive():
    rng = np.random.default_rng(3)
    n, w = 120, 25
    C = np.where(rng.random((n, w + 1)) < 0.5, -rng.random((n, w + 1)), 0.0)
    C[:, 1:][np.arange(n)[:, None] + np.arange(1, w + 1) >= n] = 0.0
    C[:, 0] = 2.0 * w
    for col in (0, 1, 57, n - 1):
        Cb = C.copy()
        Cb[col, 0] = -1.0
        with np.errstate(invalid="ignore"):
            _, _, badw = factor_window(Cb, w)
            _, badg = factor_global(Cb, w)
        assert badw == badg == col + 1
