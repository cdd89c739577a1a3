"""Pins of the oracle's projection (reference vfx/sim3d/pressure.py) by exact linear algebra on
small grids:

  * K, assembled column by column from the oracle's matvec, equals D D^T where D is the
    face-averaged divergence assembled independently face by face from the definition
    (pressure.py:3-8); K is symmetric PSD with the constant null space; diag(K) = kdiag;
  * the level-1 operators equal the component-diagonal blocks of T K T^T with T the orthonormal
    2x2x2 Hadamard transform (pressure.py:10-15, :220-246), and every coarse level equals
    (1/16) P^T A P with P the constant prolongation (:249-268, GALERKIN_SCALE :783);
  * the preconditioner is symmetric positive definite and reduces the error in the energy norm;
  * the FP64 PCG projection equals an exact sparse solve (scipy), leaves d(u) = 0, and removes
    a pure-gradient flow (tests3d/test_sim3d_layer1.py:148-165).
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from simhelp import sealed_cell_solid, smooth_noise, sphere_solid
from vfx.core.fields3d import Forces3D
from vfx.sim3d import SolverConfig3D
from vfx.sim3d.solver import pack_params


def _assemble_K(fm):
    N = fm.shape[0]
    n = N ** 3
    cols = []
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        cols.append(oracle.kmul(fm, e.reshape(N, N, N), prec="f64").ravel())
    return np.stack(cols, axis=1)


def _assemble_D(solid):
    """d(u)_i = sum over open faces (i, j) along axis a of s (u_i^a + u_j^a), s = +1 for the
    +a face and -1 for the -a face; faces are open when both cells are fluid (walls closed)."""
    N = solid.shape[0]
    n = N ** 3
    fl = solid < 0.5
    D = np.zeros((n, 3 * n))
    idx = np.arange(n).reshape(N, N, N)
    for ax, (dz, dy, dx) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
        for z in range(N):
            for y in range(N):
                for x in range(N):
                    zz, yy, xx = z + dz, y + dy, x + dx
                    if zz >= N or yy >= N or xx >= N or not (fl[z, y, x] and fl[zz, yy, xx]):
                        continue
                    i, j = idx[z, y, x], idx[zz, yy, xx]
                    for row, sgn in ((i, 1.0), (j, -1.0)):
                        D[row, ax * n + i] += sgn
                        D[row, ax * n + j] += sgn
    return D


@pytest.mark.parametrize("N,r", [(6, 0.0), (8, 0.3)])
def test_K_equals_D_DT(N, r):
    solid = sphere_solid(N, r=r) if r > 0 else np.zeros((N,) * 3, np.float32)
    fm = oracle.mask(solid)
    K = _assemble_K(fm)
    D = _assemble_D(solid)
    assert np.array_equal(K, D @ D.T)
    assert np.array_equal(K, K.T)
    w = np.linalg.eigvalsh(K)
    assert w.min() > -1e-9
    fluid = (fm.ravel() != 0)
    assert np.abs(K @ fluid.astype(float)).max() < 1e-12     # constants are in the null space
    # div2 is D applied to a velocity
    u = smooth_noise(N, 3, 4).astype(np.float64)
    np.testing.assert_allclose(oracle.div2(fm, u, prec="f64").ravel(), D @ u.ravel(), atol=1e-12)
    # apply: u - g(phi) with g = -D^T (per component)
    phi = smooth_noise(N, 1, 5)[0].astype(np.float64) * fluid.reshape(N, N, N)
    got = oracle.apply(fm, phi, u, prec="f64")
    ref = (u.ravel() + D.T @ phi.ravel()).reshape(3, N, N, N) * fluid.reshape(N, N, N)
    np.testing.assert_allclose(got, ref, atol=1e-12)


def _hadamard_T(N):
    """Orthonormal block transform: row (c, B) = t_{c,B} over the fine grid."""
    m = N // 2
    n = N ** 3
    T = np.zeros((8 * m ** 3, n))
    hs = 1.0 / np.sqrt(8.0)
    for c in range(8):
        for K in range(m):
            for J in range(m):
                for I in range(m):
                    row = c * m ** 3 + (K * m + J) * m + I
                    for dz in (0, 1):
                        for dy in (0, 1):
                            for dx in (0, 1):
                                e = (c & 1) * dx + ((c >> 1) & 1) * dy + ((c >> 2) & 1) * dz
                                col = ((2 * K + dz) * N + 2 * J + dy) * N + 2 * I + dx
                                T[row, col] = -hs if e & 1 else hs
    return T


def _level_matrix(diag, off):
    """Per-component 7-point operator from (diag (8,m,m,m), off (24,m,m,m))."""
    m = diag.shape[1]
    n = 8 * m ** 3
    A = np.zeros((n, n))
    for c in range(8):
        for K in range(m):
            for J in range(m):
                for I in range(m):
                    i = c * m ** 3 + (K * m + J) * m + I
                    A[i, i] = diag[c, K, J, I]
                    for ax, (dk, dj, di) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
                        K2, J2, I2 = K + dk, J + dj, I + di
                        if K2 < m and J2 < m and I2 < m:
                            j = c * m ** 3 + (K2 * m + J2) * m + I2
                            A[i, j] = A[j, i] = off[ax * 8 + c, K, J, I]
    return A


@pytest.mark.parametrize("N,r", [(8, 0.0), (8, 0.3), (12, 0.25)])
def test_level_operators_are_THT_blocks_and_galerkin(N, r):
    solid = sphere_solid(N, r=r) if r > 0 else np.zeros((N,) * 3, np.float32)
    fm = oracle.mask(solid)
    K = _assemble_K(fm)
    T = _hadamard_T(N)
    assert np.allclose(T @ T.T, np.eye(T.shape[0]))
    A1 = T @ K @ T.T
    ops = oracle.build_ops(fm, prec="f64")
    d1, o1 = ops[0]
    L1 = _level_matrix(d1, o1)
    m = N // 2
    # same entries on the component diagonal (c = c') for the diagonal and nearest-neighbour blocks
    comp = np.repeat(np.arange(8), m ** 3)
    same = comp[:, None] == comp[None, :]
    mask = same & (L1 != 0)
    np.testing.assert_allclose(L1[mask], A1[mask], atol=1e-12)
    # and the component-diagonal part of A1 has no other couplings
    extra = same & (L1 == 0) & (np.abs(A1) > 1e-12)
    assert not extra.any()
    # coarse levels: (1/16) P^T A P per component, P = constant prolongation
    for (df, of), (dc, oc) in zip(ops[:-1], ops[1:]):
        mf, mc = df.shape[1], dc.shape[1]
        Af = _level_matrix(df, of)
        P = np.zeros((8 * mf ** 3, 8 * mc ** 3))
        for c in range(8):
            for K_ in range(mf):
                for J_ in range(mf):
                    for I_ in range(mf):
                        P[c * mf ** 3 + (K_ * mf + J_) * mf + I_,
                          c * mc ** 3 + ((K_ // 2) * mc + J_ // 2) * mc + I_ // 2] = 1.0
        G = P.T @ Af @ P / 16.0
        np.testing.assert_allclose(_level_matrix(dc, oc), G, atol=1e-12)


@pytest.mark.parametrize("N,r", [(8, 0.3), (12, 0.25), (16, 0.2)])
def test_preconditioner_spd_and_reducing(N, r):
    solid = sphere_solid(N, r=r)
    fm = oracle.mask(solid)
    fl = (fm != 0).astype(np.float64)
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N,) * 3) * fl
    y = rng.standard_normal((N,) * 3) * fl
    Mx, _ = oracle.precond_apply(fm, x, prec="f64")
    My, _ = oracle.precond_apply(fm, y, prec="f64")
    asym = abs((x * My).sum() - (Mx * y).sum()) / abs((x * My).sum())
    if oracle.level_sizes(N)[-1] == 1:
        # V(1,1) with red-black pre- and post-sweeps in mirrored order and an exact coarsest
        # solve: a symmetric operator
        assert asym <= 1e-10
    else:
        # odd coarsest level (pressure.py:916-922): 2m forward red-black sweep pairs are not a
        # symmetric coarse solve, so M^-1 is only nearly symmetric (as in the reference)
        assert asym <= 0.05
    assert (x * Mx).sum() > 0 and (y * My).sum() > 0
    # one preconditioned Richardson step reduces a smooth error in the K-norm
    e = smooth_noise(N, 1, 9)[0].astype(np.float64) * fl
    e -= e[fl > 0].mean() * fl
    Ke = oracle.kmul(fm, e, prec="f64")
    MKe, _ = oracle.precond_apply(fm, Ke, prec="f64")
    e2 = e - MKe
    en = lambda v: float((v * oracle.kmul(fm, v, prec="f64")).sum())
    assert en(e2) < 0.5 * en(e)
    # the FP32 instantiation is the same operator to FP32 round-off
    Mx32, _ = oracle.precond_apply(fm, x.astype(np.float32))
    assert np.abs(Mx32 - Mx).max() <= 1e-5 * np.abs(Mx).max()


def test_round16_is_the_library_conversion():
    """The oracle's binary16 / bfloat16 rounding equals torch's float -> half / bfloat16 conversion
    (round to nearest, ties to even) value for value: normal and subnormal binary16 ranges,
    overflow to infinity, exact ties, and bfloat16's 7-bit fraction."""
    import torch
    rng = np.random.default_rng(7)
    vals = np.concatenate([
        rng.standard_normal(4000) * 10.0 ** rng.uniform(-9, 5, 4000),
        (rng.integers(1, 2048, 500) + 0.5) * 2.0 ** -10,          # exact binary16 ties near 1
        (rng.integers(1, 1024, 500) + 0.5) * 2.0 ** -24,          # ties in the subnormal range
        (rng.integers(1, 128, 500) + 0.5) * 2.0 ** -7,            # exact bfloat16 ties near 1
        [65504.0, 65519.0, 65520.0, 1e6, -65520.0, 6.1035e-05, 6.10351e-05, 0.0, -0.0],
    ]).astype(np.float32)
    t = torch.from_numpy(vals)
    for kind, conv in (("fp16", torch.float16), ("bf16", torch.bfloat16)):
        got = oracle.round16(vals, kind)
        ref = t.to(conv).float().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), kind


def test_precond16_fp32_storage_is_the_fp32_path():
    """With FP32 storage the 16-bit-capable preconditioner is the pinned FP32 one, bit for bit."""
    N = 16
    fm = oracle.mask(sphere_solid(N, r=0.25))
    r = smooth_noise(N, 1, 3)[0] * (fm != 0)
    a, _ = oracle.precond_apply(fm, r)
    b, _ = oracle.precond_apply16(fm, r, "fp32")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("storage,tol", [("fp16", 4e-3), ("bf16", 4e-2)])
def test_precond16_scale_and_closeness(storage, tol):
    """16-bit level storage: (1) the level-1 right-hand side is scaled by a power of two taken from
    max|r| and the scale is undone by the inverse transform, so M^-1 (2^k r) = 2^k M^-1 r exactly
    even for k that would underflow or overflow the 16-bit range without it; (2) the operator
    stays within the storage precision of the FP32 one; (3) it still reduces a smooth error in the
    K-norm."""
    N = 16
    fm = oracle.mask(sphere_solid(N, r=0.25))
    fl = (fm != 0).astype(np.float32)
    r = (smooth_noise(N, 1, 5)[0] * fl).astype(np.float32)
    z, _ = oracle.precond_apply16(fm, r, storage)
    for k in (-40, 30):
        zk, _ = oracle.precond_apply16(fm, (r * np.float32(2.0 ** k)).astype(np.float32), storage)
        assert np.array_equal(zk, (z * np.float32(2.0 ** k)).astype(np.float32)), k
    z32, _ = oracle.precond_apply(fm, r)
    err = np.abs(z - z32).max() / np.abs(z32).max()
    assert 0 < err <= tol, err
    e = smooth_noise(N, 1, 9)[0].astype(np.float64) * fl
    e -= e[fl > 0].mean() * fl
    Ke = oracle.kmul(fm, e, prec="f64").astype(np.float32)
    MKe, _ = oracle.precond_apply16(fm, Ke, storage)
    e2 = e - MKe
    en = lambda v: float((v * oracle.kmul(fm, v, prec="f64")).sum())
    assert en(e2) < 0.5 * en(e)


def _exact_projection(solid, u):
    """u - g(phi) with K phi = -d(u) solved exactly (sparse LU on the fluid cells, mean fixed)."""
    N = solid.shape[0]
    fm = oracle.mask(solid)
    K = sp.csr_matrix(_assemble_K(fm))
    fl = fm.ravel() != 0
    b = -oracle.div2(fm, u, prec="f64").ravel()
    idx = np.flatnonzero(fl)
    Kf = K[idx][:, idx].tolil()
    bf = b[idx].copy()
    # pin one cell of the main component to fix the constant, and every sealed cell (a zero row
    # and column of K: a component of its own)
    for j in [0] + list(np.flatnonzero(Kf.diagonal() == 0)):
        Kf[j, :] = 0
        Kf[j, j] = 1.0
        bf[j] = 0.0
    phi = np.zeros(N ** 3)
    phi[idx] = spla.spsolve(Kf.tocsr(), bf)
    return oracle.apply(fm, phi.reshape(N, N, N), u, prec="f64"), phi


@pytest.mark.parametrize("N,r", [(8, 0.3), (12, 0.25)])
def test_pcg_equals_exact_solve(N, r):
    solid = sphere_solid(N, r=r)
    u = (smooth_noise(N, 3, 2, amp=2.0) + np.array([0.8, -0.3, 0.4], np.float32)[:, None, None, None])
    u = u.astype(np.float64) * (solid < 0.5)
    ref, _ = _exact_projection(solid, u)
    data = np.zeros((6, N, N, N))
    data[1:4] = u
    cfg = SolverConfig3D(pressure_rel_tol=1e-12, pressure_max_iters=500, pressure_warm_start=False)
    it, st = oracle.project(data, solid, pack_params(cfg, N, Forces3D(), 0.0), np.zeros((N,) * 3), prec="f64")
    assert st["r_norm"] <= 1e-12 * st["b_norm"] and 0 < it < 100
    np.testing.assert_allclose(data[1:4], ref, atol=1e-9 * np.abs(u).max())
    fm = oracle.mask(solid)
    assert np.abs(oracle.div2(fm, data[1:4], prec="f64")).max() < 1e-9


def test_sealed_fluid_cell():
    """A fluid cell sealed in by solids on all six sides has diag(K) = 0 and a zero row and column
    of K.  Its Jacobi terms are 0, so the preconditioner stays finite, symmetric and positive on
    the rest (the reference divides 0 / 0 there, pressure.py:671 and :702, and its PCG stops at
    iteration 0 on the NaN), and the projection equals the exact solve with the sealed cell left
    alone."""
    N, c = 12, (3, 4, 3)
    solid = sealed_cell_solid(N, c)
    fm = oracle.mask(solid)
    assert fm[c] == 64 and oracle.kmul(fm, np.ones((N,) * 3), prec="f64")[c] == 0.0
    fl = (fm != 0).astype(np.float64)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((N,) * 3) * fl
    y = rng.standard_normal((N,) * 3) * fl
    Mx, _ = oracle.precond_apply(fm, x, prec="f64")
    My, _ = oracle.precond_apply(fm, y, prec="f64")
    assert np.isfinite(Mx).all() and Mx[c] == 0.0
    assert (x * Mx).sum() > 0 and (y * My).sum() > 0
    assert abs((x * My).sum() - (Mx * y).sum()) <= 0.05 * abs((x * My).sum())
    u = (smooth_noise(N, 3, 4, amp=2.0) + np.array([0.5, 0.2, -0.4], np.float32)[:, None, None, None])
    u = u.astype(np.float64) * (solid < 0.5)
    ref, _ = _exact_projection(solid, u)
    data = np.zeros((6, N, N, N))
    data[1:4] = u
    cfg = SolverConfig3D(pressure_rel_tol=1e-12, pressure_max_iters=500, pressure_warm_start=False)
    it, st = oracle.project(data, solid, pack_params(cfg, N, Forces3D(), 0.0), np.zeros((N,) * 3), prec="f64")
    assert 0 < it < 100 and st["r_norm"] <= 1e-12 * st["b_norm"]
    np.testing.assert_allclose(data[1:4], ref, atol=1e-9 * np.abs(u).max())
    assert np.array_equal(data[1:4, c[0], c[1], c[2]], u[:, c[0], c[1], c[2]])   # g = 0 there
    # the FP32 solver converges too
    d32 = np.zeros((6, N, N, N), np.float32)
    d32[1:4] = u
    cfg32 = SolverConfig3D(pressure_rel_tol=1e-5, pressure_max_iters=200, pressure_warm_start=False)
    it32, st32 = oracle.project(d32, solid, pack_params(cfg32, N, Forces3D(), 0.0), np.zeros((N,) * 3, np.float32))
    assert 0 < it32 < 40 and st32["r_norm"] <= 1e-5 * st32["b_norm"] and np.isfinite(d32).all()


def test_pure_gradient_flow_is_removed():
    """u = grad(x + 0.5 y - 0.3 z) in a closed box projects to ~0 (test_sim3d_layer1.py:148-165)."""
    N = 16
    data = np.zeros((6, N, N, N), np.float32)
    data[1], data[2], data[3] = 1.0, 0.5, -0.3
    cfg = SolverConfig3D(pressure_tol=1e-7, pressure_max_iters=5000, pressure_warm_start=False)
    solid = np.zeros((N,) * 3, np.float32)
    it, _ = oracle.project(data, solid, pack_params(cfg, N, Forces3D(), 0.0), np.zeros((N,) * 3, np.float32))
    u = data[1:4]
    assert np.abs(u).max() < 1e-2
    assert max(np.abs(u[0].mean(axis=(0, 1))).max(), np.abs(u[1].mean(axis=(0, 2))).max(),
               np.abs(u[2].mean(axis=(1, 2))).max()) < 1e-2


def test_reference_stop_test_and_warm_start():
    """Normalised max-residual stop (k_fin_maxu / k_fin_rz, pressure.py:730-771); with a warm
    start from the converged phi the next solve needs no iteration."""
    N = 16
    solid = sphere_solid(N, r=0.2)
    u = smooth_noise(N, 3, 11, amp=2.0) * (solid < 0.5)
    data = np.zeros((6, N, N, N), np.float32)
    data[1:4] = u
    cfg = SolverConfig3D(pressure_tol=2e-5, pressure_max_iters=40)
    prm = pack_params(cfg, N, Forces3D(), 0.0)
    phi = np.zeros((N,) * 3, np.float32)
    it, st = oracle.project(data, solid, prm, phi)
    fm = oracle.mask(solid)
    maxr = np.abs(oracle.div2(fm, data[1:4])).max()
    assert 0 < it < 40
    assert maxr <= 2 * 2e-5 * st["max_u"] * 1.01
    data2 = np.zeros_like(data)
    data2[1:4] = u
    it2, _ = oracle.project(data2, solid, prm, phi)
    assert it2 == 0
    np.testing.assert_allclose(data2[1:4], data[1:4], atol=1e-5 * np.abs(u).max())


# ---------------------------------------------------------------- the fused projection
def test_fused_weights_are_chebyshev():
    """The sweep weights are the reciprocals of the degree-n Chebyshev nodes on [0.5, 2]
    (numpy's chebroots), so the sweep polynomial prod(1 - w lambda) has the minimax value
    1 / T_n(5/3) on that interval (pressure_fused.py:348-352)."""
    for n in (1, 2, 4, 6):
        w = np.array(oracle.fused_weights(n))
        t = np.polynomial.chebyshev.chebroots([0] * n + [1])
        nodes = np.sort(1.25 + 0.75 * t)
        np.testing.assert_allclose(np.sort(1.0 / w), nodes, rtol=1e-12)
        lam = np.linspace(0.5, 2.0, 20001)
        p = np.prod(1.0 - w[:, None] * lam[None, :], axis=0)
        tn = np.cosh(n * np.arccosh(5.0 / 3.0))
        assert abs(np.abs(p).max() - 1.0 / tn) < 1e-6
    assert oracle.fused_weights(3, "jacobi") == [0.8, 0.8, 0.8]


def _kdiag_np(fm):
    s = np.zeros(fm.shape)
    for ax in range(3):
        oR = (fm >> (2 * ax)) & 1
        oL = (fm >> (2 * ax + 1)) & 1
        s += (oL - oR) ** 2 + oL + oR
    return s


def test_fused_single_sweep_is_weighted_jacobi():
    """cycles = 0, one sweep from a warm start: phi_1 = phi_0 + w D^-1 (b - K phi_0) with
    b = -d(u*), then u -= g(phi_1) (assembled here from the oracle's operators)."""
    N = 12
    solid = sphere_solid(N, r=0.25)
    fm = oracle.mask(solid)
    u = (smooth_noise(N, 3, 6, amp=1.0) * (solid < 0.5)).astype(np.float64)
    phi0 = smooth_noise(N, 1, 7)[0].astype(np.float64) * (fm != 0)
    cfg = SolverConfig3D(pressure_tol=0.0, pressure_warm_start=True)
    data = np.zeros((6, N, N, N))
    data[1:4] = u
    phi = phi0.copy()
    it, _ = oracle.project_fused(data, solid, pack_params(cfg, N, Forces3D(), 0.0), phi, 0, 1, prec="f64")
    assert it == 0
    w = oracle.fused_weights(1)[0]
    b = -oracle.div2(fm, u, prec="f64")
    w32 = float(np.float32(w))
    ref = np.where(fm != 0, phi0 + w32 * (b - oracle.kmul(fm, phi0, prec="f64")) / np.maximum(_kdiag_np(fm), 1), 0.0)
    np.testing.assert_allclose(phi, ref, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(data[1:4], oracle.apply(fm, ref, u, prec="f64"), rtol=1e-13, atol=1e-14)


def test_fused_adaptive_converges_to_exact_solve():
    """With the adaptive tail (tolerance 1e-10, at most 200 cycles) the stationary multigrid
    iteration reaches the exact projection; with a fixed 2 cycles it is already close."""
    N = 12
    solid = sphere_solid(N, r=0.25)
    u = (smooth_noise(N, 3, 2, amp=2.0) + np.array([0.8, -0.3, 0.4], np.float32)[:, None, None, None])
    u = u.astype(np.float64) * (solid < 0.5)
    ref, _ = _exact_projection(solid, u)
    errs = {}
    for name, tol, maxit in (("fixed", 0.0, 24), ("adaptive", 1e-10, 200)):
        cfg = SolverConfig3D(pressure_tol=tol, pressure_max_iters=maxit, pressure_warm_start=False)
        data = np.zeros((6, N, N, N))
        data[1:4] = u
        it, st = oracle.project_fused(data, solid, pack_params(cfg, N, Forces3D(), 0.0), np.zeros((N,) * 3), 2, 4,
                                      prec="f64")
        errs[name] = (it, float(np.abs(data[1:4] - ref).max() / np.abs(u).max()), st)
    assert errs["fixed"][0] == 2 and errs["fixed"][1] < 0.05
    it, e, st = errs["adaptive"]
    assert 2 < it < 200 and 0.5 * st["max_r"] < 1e-10 * st["max_u"]
    assert e < 1e-8


def test_fused_sealed_cell_stays_finite():
    """A sealed fluid cell (kdiag = 0) keeps its phi in the sweeps: the fused projection stays
    finite and matches the exact solve on the rest."""
    N, c = 12, (3, 4, 3)
    solid = sealed_cell_solid(N, c)
    u = (smooth_noise(N, 3, 4, amp=2.0) + np.array([0.5, 0.2, -0.4], np.float32)[:, None, None, None])
    u = u.astype(np.float64) * (solid < 0.5)
    ref, _ = _exact_projection(solid, u)
    cfg = SolverConfig3D(pressure_tol=1e-10, pressure_max_iters=300, pressure_warm_start=False)
    data = np.zeros((6, N, N, N))
    data[1:4] = u
    phi = np.zeros((N,) * 3)
    it, _ = oracle.project_fused(data, solid, pack_params(cfg, N, Forces3D(), 0.0), phi, 2, 4, prec="f64")
    assert np.isfinite(data).all() and np.isfinite(phi).all()
    assert np.abs(data[1:4] - ref).max() < 1e-7 * np.abs(u).max()
