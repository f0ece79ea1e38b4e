"""FP64 CPU restatement of QUMASUN's block LOBPCG (LOBPCG::Eigen_z_multi_mpi).

TEST INFRASTRUCTURE ONLY (checker for the GPU solver; never shipped or timed
as the product).  Follows lobpcg_z_multi_mpi.h:1732-2852:

* x is scaled by sqrt(dV) on entry so that plain inner products are
  normalised (1851), and unscaled / renormalised on exit (2796-2800).
* eps_n = Re <x_n|A x_n>, r_n = A x_n - eps_n x_n (1873-1897).
* P = 0 on the first call (2052-2056); otherwise A P is recomputed (2058-2065).
* per iteration: A R (2084); Rayleigh-Ritz on [X R P] (or [X R] while P is
  empty, 2124-2288) through the generalized eigenproblem H c = l S c (the
  lapack_ZHEGVD(_emu) step, w_zhegvd.h:227-272; here scipy.linalg.eigh, which
  is LAPACK's *sygvd/*hegvd);
  X' = X Cx + R Cr + P Cp, P' = R Cr + P Cp, AX', AP' likewise (NextVector_z2, 1515-1598);
  loss of normalisation |1 - |x'|^2| > 1e-3 -> P = R, renormalise the old X,
  leave the loop (2611-2666); otherwise R' = AX' - X' L (2750).
* The subspace problem is solved through the Cholesky factor of S (the
  lapack_ZHEGVD_emu route, w_zhegvd.h:227-272).  When S is not numerically
  positive definite, or its squared Cholesky pivots span less than `cond`, the
  step is redone on [X R] (P dropped) and then on the S eigenbasis with
  near-null directions dropped (rr_solve, subspace_eig).  The same happens when
  the N lowest Ritz vectors of the Cholesky route are not S-orthonormal to
  RITZ_ORTH_TOL (this build's guard; the reference's zhegvd has none).
* Gram blocks: every block computed directly on iteration 0 and every 6th
  iteration (S_MAT_REST_LIMIT = 6, 1808, 2093); in between the X/P blocks come
  from the subspace recurrence (NextMatrix_z_blasmp, 1104-1436:
  X'^H X' = Cx^H S Cx, X'^H P' = Cx^H S Cp', P'^H P' = Cp'^H S Cp', and the same with H).
  This build computes the R blocks (X^H R, R^H R, P^H R and their H versions)
  directly every iteration, where the reference takes X^H R and R^H P from the
  recurrence (DIRECT_S_MATRIX_XR_RP, 58).
* Before A R, the residual block is replaced by an orthonormal basis of its
  component orthogonal to X (Cholesky QR, dropping near-dependent directions);
  P columns are normalised (the reference's EVERY_NORMALIZE_P option,
  2690-2707).  Both are changes of basis of span{X, R, P}, so the iterates are
  unchanged in exact arithmetic, and S stays well conditioned.
* Locking (residual_threshold, num_fix; 1962-1986): with a threshold set,
  num_fix (kept on the instance between calls, like the caller-owned
  m_num_fix_state_sk, and reset by the caller) becomes the first band at or
  above its value on entry whose |r|^2 exceeds threshold^2, at most N/2.  Only
  the bands above it iterate; their R every iteration, and P when num_fix
  changed, lose their components along the fixed bands first
  (VectorOrthogonalization, 1637-1723: V - X_f (X_f^H V)).  The fixed bands
  are returned unscaled and their eigenvalue entries are left as passed
  (2651-2652, 2796-2797, 2846-2848).  Without a threshold num_fix stays 0.

Vectors are complex internally; the SoA real layout (re, im rows per band, or
one row per band for real Gamma-point bands) is converted at the boundary.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

LIMIT_NORM_X = 1.0e-3   # LIMIT_Z_NORM_X_FOR_RESET_MATRIX (lobpcg_z_multi_mpi.h:101)
S_MAT_REST_LIMIT = 6    # lobpcg_z_multi_mpi.h:1808


def soa_to_cplx(a, ncomp):
    if ncomp == 1:
        return np.array(a, dtype=np.float64)
    return a[0::2] + 1j * a[1::2]


def cplx_to_soa(z, ncomp):
    if ncomp == 1:
        return np.ascontiguousarray(z.real if np.iscomplexobj(z) else z)
    out = np.empty((2 * z.shape[0], z.shape[1]))
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


SVQB_DROP = 1.0e-12     # relative eigenvalue of S below which a direction is dropped
RITZ_ORTH_TOL = 1.0e-4  # max |1 - c_i^H S c_i| of accepted Cholesky-route Ritz vectors


def subspace_eig(Hm, S, pen=1.0e30):
    """H c = l S c.  Cholesky-based generalized solve (LAPACK *sygvd/*hegvd, the
    reference's lapack_ZHEGVD path); when S is not numerically positive
    definite (R and P nearly dependent near convergence, where the reference's
    LAPACK call would report info > 0), fall back to the S eigenbasis with the
    near-null directions dropped (SVQB): B = V s^-1/2, eigh(B^H H B), C = B y.
    Dropped directions get the diagonal `pen` (above the N lowest Ritz values,
    see dropped_penalty) so they sort above them."""
    try:
        return scipy.linalg.eigh(Hm, S, driver="gvd")
    except np.linalg.LinAlgError:
        s, V = scipy.linalg.eigh(S, driver="evd")
        keep = s > SVQB_DROP * s.max()
        w = np.where(keep, 1.0 / np.sqrt(np.where(keep, s, 1.0)), 0.0)
        B = V * w[None, :]
        Hr = np.conj(B).T @ Hm @ B
        Hr = 0.5 * (Hr + np.conj(Hr).T)
        Hr[np.diag_indices_from(Hr)] += np.where(keep, 0.0, pen)
        lam, Y = scipy.linalg.eigh(Hr, driver="evd")
        return lam, B @ Y


def orthonormalise_r(x, r, drop):
    """R' = [X Rn] W with Rn = R normalised per column: R orthogonalised against X
    (X^H X = I) and orthonormalised through the Cholesky factor of
    M = Rn^H Rn - (X^H Rn)^H (X^H Rn); when M's squared pivots span less than
    `drop`, through its eigenbasis with near-null directions dropped (zero rows).
    Returns (R', keep mask)."""
    N = x.shape[0]
    rn = np.sqrt(np.sum(np.abs(r) ** 2, axis=1))
    rn[rn == 0] = 1.0
    r = r / rn[:, None]
    XR = np.conj(x) @ r.T
    M = np.conj(r) @ r.T - np.conj(XR).T @ XR
    M = 0.5 * (M + np.conj(M).T)
    try:
        U = scipy.linalg.cholesky(M, lower=False)
        d = np.abs(np.diag(U)) ** 2
        if d.min() < drop * d.max():
            raise np.linalg.LinAlgError
        Uinv = scipy.linalg.solve_triangular(U, np.eye(N), lower=False)
        W = np.concatenate([-XR @ Uinv, Uinv], axis=0)
        keep = np.ones(N, dtype=bool)
    except np.linalg.LinAlgError:
        sv, V = scipy.linalg.eigh(M, driver="evd")
        keep = sv > drop * sv.max()
        Vs = V * np.where(keep, 1.0 / np.sqrt(np.where(keep, sv, 1.0)), 0.0)[None, :]
        W = np.concatenate([-XR @ Vs, Vs], axis=0)
    return W[:N].T @ x + W[N:].T @ r, keep


def dropped_penalty(Hm, N):
    """Diagonal given to dropped directions: twice a Gershgorin bound of the X
    block's H plus one, above its largest Ritz value and so above the N-th Ritz
    value of any subspace containing X (a fixed 1e30 would put eps * 1e30 of
    rounding on every eigenvalue of a dense eigensolver)."""
    return 2.0 * np.abs(Hm[:N, :N]).sum(axis=1).max() + 1.0


def rr_solve(Hm, S, N, have_p, cond):
    """Rayleigh-Ritz step as the GPU driver takes it: Cholesky of S with a
    squared-pivot ratio >= cond, else retry on [X R] (P dropped), else the
    S-eigenbasis solve.  Returns (lam, C, used_p)."""
    sizes = [3 * N, 2 * N] if have_p else [2 * N]
    for n3 in sizes:
        Ss, Hs = S[:n3, :n3], Hm[:n3, :n3]
        try:
            U = scipy.linalg.cholesky(Ss, lower=False)
        except np.linalg.LinAlgError:
            continue
        d = np.abs(np.diag(U)) ** 2
        if d.min() < cond * d.max():
            continue
        lam, C = scipy.linalg.eigh(Hs, Ss, driver="gvd")
        # the N lowest Ritz vectors must be S-orthonormal to 1e-4 (an S too
        # ill-conditioned for its Cholesky factor breaks this first)
        CN = C[:, :N]
        dev = np.abs(1.0 - np.real(np.einsum("ji,jk,ki->i", np.conj(CN), Ss, CN)))
        if not np.all(dev <= RITZ_ORTH_TOL):
            continue
        return lam, C, n3 == 3 * N
    n3 = 2 * N
    lam, C = subspace_eig(Hm[:n3, :n3], S[:n3, :n3], dropped_penalty(Hm, N))
    return lam, C, False


class LOBPCGRef:
    """Keeps P between calls like the reference's `keep` block."""

    def __init__(self, apply_soa, ncomp, dV, cond=1.0e-14, rdrop=1.0e-13, residual_threshold=None):
        self.apply_soa = apply_soa
        self.ncomp = ncomp
        self.dV = dV
        self.cond = cond    # minimum squared Cholesky pivot ratio of S
        self.rdrop = rdrop  # R directions kept above this pivot / eigenvalue ratio
        self.residual_threshold = residual_threshold  # None: no locking
        self.limit_norm_x = LIMIT_NORM_X
        self.num_fix = 0    # fixed lowest bands, kept between calls
        self.P = None       # all N bands (rows below num_fix unused)

    def _A(self, z):
        return soa_to_cplx(self.apply_soa(cplx_to_soa(z, self.ncomp)), self.ncomp)

    @staticmethod
    def _gram(a, b):
        """a^H b with bands as rows: (Na x Ng) . (Nb x Ng)^T."""
        return np.conj(a) @ b.T

    def _orthogonalise_fixed(self, xf, v):
        """VectorOrthogonalization (lobpcg_z_multi_mpi.h:1637-1723): V - X_f (X_f^H V),
        bands as rows."""
        if xf.shape[0] == 0:
            return v
        return v - self._gram(xf, v).T @ xf

    def solve(self, X_soa, iter_max, first_time=True, return_history=False, eigen_values=None):
        """Returns (X, eigenvalues, |r|^2, iterations[, history]) over all N bands;
        entries of the bands fixed on return are taken from `eigen_values` (NaN
        if None) and NaN in |r|^2."""
        nc = self.ncomp
        dt = np.complex128 if nc == 2 else np.float64
        xall = soa_to_cplx(X_soa, nc).astype(dt) * np.sqrt(self.dV)
        Nall = xall.shape[0]
        lock = self.residual_threshold is not None
        f0 = self.num_fix if lock else 0
        assert 0 <= f0 <= Nall // 2
        x = xall[f0:]
        Ax = self._A(x)
        eps = np.real(np.sum(np.conj(x) * Ax, axis=1))
        r = Ax - eps[:, None] * x
        rnorm2 = np.sum(np.abs(r) ** 2, axis=1)
        f = f0
        if lock:
            # first band at or above num_fix whose |r|^2 exceeds threshold^2, at most N/2 (1962-1986)
            thr2 = self.residual_threshold ** 2
            pos = next((k for k in range(f0, Nall) if thr2 < rnorm2[k - f0]), Nall)
            f = min(pos, Nall // 2)
            self.num_fix = f
        d = f - f0
        x, Ax, r, eps, rnorm2 = x[d:], Ax[d:], r[d:], eps[d:], rnorm2[d:]
        xf = xall[:f]
        N = Nall - f
        ev_full = np.full(Nall, np.nan) if eigen_values is None else np.array(eigen_values, dtype=np.float64)
        rn_full = np.full(Nall, np.nan)

        def finish(xa, ev_a, rn_a, it_done, history):
            # fixed bands unscaled (MulC_bundle(num_fix, x, 1/sqrt(dV)), 2651-2652 / 2796-2797)
            out = cplx_to_soa(np.concatenate([xf / np.sqrt(self.dV), xa]), nc)
            ev_full[f:] = ev_a
            rn_full[f:] = rn_a
            res = (out, ev_full, rn_full, it_done)
            return res + (history,) if return_history else res

        def keep_p(pa):
            self.P = np.concatenate([np.zeros((f, pa.shape[1]), dtype=dt), pa])

        if iter_max < 1:
            # no iteration: undo the scaling and leave (1992-2004)
            return finish(x / np.sqrt(self.dV), eps, rnorm2, 0, [])
        if first_time or self.P is None:
            p = np.zeros_like(x)
            Ap = np.zeros_like(x)
            have_p = False
        else:
            p = self.P[f:].astype(dt)
            if f != f0:  # the active set changed: P loses its fixed-band components (2058-2063)
                p = self._orthogonalise_fixed(xf, p)
            Ap = self._A(p)
            have_p = True
        blocks = {}
        history = []
        rnorm2 = np.sum(np.abs(r) ** 2, axis=1)
        it_done = 0
        for it in range(iter_max):
            it_done = it + 1
            # R loses its fixed-band components (2080), then R <- orthonormal basis of
            # (1 - X X^H) R (exact-arithmetic invariant), then A R
            r = self._orthogonalise_fixed(xf, r)
            r, rkeep = orthonormalise_r(x, r, self.rdrop)
            Ar = self._A(r)
            G = self._gram
            recompute = (it % S_MAT_REST_LIMIT == 0)
            if recompute:
                blocks["XX"], blocks["XAX"] = G(x, x), G(x, Ax)
                if have_p:
                    blocks["XP"], blocks["PP"] = G(x, p), G(p, p)
                    blocks["XAP"], blocks["PAP"] = G(x, Ap), G(p, Ap)
            blocks["XR"], blocks["RR"], blocks["XAR"], blocks["RAR"] = G(x, r), G(r, r), G(x, Ar), G(r, Ar)
            if have_p:
                # R^H P and R^H A P taken as (P^H R)^H and (P^H A R)^H, as the GPU Gram does
                blocks["RP"], blocks["RAP"] = G(p, r).conj().T, G(p, Ar).conj().T
            nb = 3 if have_p else 2
            S = np.zeros((nb * N, nb * N), dtype=dt)
            Hm = np.zeros((nb * N, nb * N), dtype=dt)
            b = blocks
            S[:N, :N], S[:N, N:2 * N], S[N:2 * N, N:2 * N] = b["XX"], b["XR"], b["RR"]
            Hm[:N, :N], Hm[:N, N:2 * N], Hm[N:2 * N, N:2 * N] = b["XAX"], b["XAR"], b["RAR"]
            if have_p:
                S[:N, 2 * N:], S[N:2 * N, 2 * N:], S[2 * N:, 2 * N:] = b["XP"], b["RP"], b["PP"]
                Hm[:N, 2 * N:], Hm[N:2 * N, 2 * N:], Hm[2 * N:, 2 * N:] = b["XAP"], b["RAP"], b["PAP"]
            # Hermitian from the upper triangle of blocks (Conj_U_to_L, conjugate.h)
            S = np.triu(S) + np.triu(S, 1).conj().T
            Hm = np.triu(Hm) + np.triu(Hm, 1).conj().T
            pen = dropped_penalty(Hm, N)
            for j in np.nonzero(~rkeep)[0]:  # dropped R directions: decoupled, sorted above the N lowest
                S[N + j, N + j] = 1.0
                Hm[N + j, N + j] = pen
            lam, C, used_p = rr_solve(Hm, S, N, have_p, self.cond)
            if not used_p:  # solved on [X R]: continue as if P were empty
                S, Hm = S[:2 * N, :2 * N], Hm[:2 * N, :2 * N]
                have_p = False
            lam, C = lam[:N], C[:, :N]
            Cx, Cr = C[:N], C[N:2 * N]
            Cp = C[2 * N:] if have_p else None
            # X' = X Cx + R Cr + P Cp ; P' = R Cr + P Cp  (bands are rows: Z' = C^T Z)
            pn = Cr.T @ r + (Cp.T @ p if have_p else 0)
            xn = Cx.T @ x + pn
            Apn = Cr.T @ Ar + (Cp.T @ Ap if have_p else 0)
            Axn = Cx.T @ Ax + Apn
            norms = np.sum(np.abs(xn) ** 2, axis=1)
            if np.max(np.abs(1.0 - norms)) > self.limit_norm_x:
                # rollback and escape (2611-2666): P = R and leave the loop.  The
                # reference rescales the old x by the norms of the rejected update
                # (Mul_invSqV_bundle with sum_norms, 2639), which leaves x
                # unnormalised exactly when the update lost normalisation; the old
                # x is renormalised with its own norms here.
                old = np.sum(np.abs(x) ** 2, axis=1)
                keep_p(r.copy())
                return finish(x / np.sqrt(old * self.dV)[:, None], eps, rnorm2, it_done, history)
            final = (it + 1 == iter_max)
            # recurrence for the X / P blocks of the next iteration (1104-1436)
            Cp0 = np.concatenate([np.zeros((N, N), dtype=dt), Cr] + ([Cp] if have_p else []), axis=0)
            Cb = np.concatenate([C, Cp0], axis=1)
            SS = np.conj(Cb).T @ (S @ Cb)
            HH = np.conj(Cb).T @ (Hm @ Cb)
            # normalise P' (and A P') by its norms, scaling the recurrence blocks alike
            pnorm = np.sqrt(np.real(np.diag(SS)[N:]))
            pnorm[pnorm == 0] = 1.0
            pn = pn / pnorm[:, None]
            Apn = Apn / pnorm[:, None]
            D = 1.0 / pnorm
            blocks["XX"] = SS[:N, :N]
            blocks["XP"] = SS[:N, N:] * D[None, :]
            blocks["PP"] = SS[N:, N:] * D[:, None] * D[None, :]
            blocks["XAX"] = HH[:N, :N]
            blocks["XAP"] = HH[:N, N:] * D[None, :]
            blocks["PAP"] = HH[N:, N:] * D[:, None] * D[None, :]
            x, p, Ax, Ap = xn, pn, Axn, Apn
            have_p = True
            eps = lam
            r = Ax - eps[:, None] * x
            rnorm2 = np.sum(np.abs(r) ** 2, axis=1)
            if return_history:
                history.append(lam.copy())
            if final:
                x = x / np.sqrt(norms * self.dV)[:, None]
        keep_p(p)
        return finish(x, eps, rnorm2, it_done, history)

def lobpcg_reference_variant(apply_soa, X_soa, ncomp, dV, iters):
    """The reference's own iteration, for pinning this build's variant against
    it: Ritz values after each iteration of one first-time call.

    Follows Eigen_z_multi_mpi (lobpcg_z_multi_mpi.h:1732-2852) with its
    default build options: R is used as computed (no orthonormalisation
    against X), P is not normalised (EVERY_NORMALIZE_P off, 71), and X^H R and
    R^H P come from the previous step's coefficients (DIRECT_S_MATRIX_XR_RP,
    58): with R = A X - X L and the recurrence blocks X^H X, X^H A X, X^H P,
    X^H A P, P^H P, P^H A P, X^H R = X^H A X - X^H X L and
    R^H P = X^H A P - L X^H P (A Hermitian).  R^H R, X^H A R, R^H A R and R^H A P
    are formed from the vectors (MakeMatrix_z_blas, 300-558).  Rayleigh-Ritz
    on [X R P] (or [X R] on the first iteration) by LAPACK's generalized
    eigensolver; X' = X Cx + R Cr + P Cp, P' = R Cr + P Cp (NextVector_z2,
    1515-1598).  The subspaces equal the build variant's (its R
    orthonormalisation and P scaling are changes of basis), so in exact
    arithmetic the Ritz values agree iteration by iteration."""
    nc = ncomp
    dt = np.complex128 if nc == 2 else np.float64
    A = lambda z: soa_to_cplx(apply_soa(cplx_to_soa(z, nc)), nc)  # noqa: E731
    G = lambda a, b: np.conj(a) @ b.T  # noqa: E731
    x = soa_to_cplx(X_soa, nc).astype(dt) * np.sqrt(dV)
    N = x.shape[0]
    Ax = A(x)
    lam = np.real(np.sum(np.conj(x) * Ax, axis=1))
    p = Ap = None
    rec = {}
    history = []
    for it in range(iters):
        r = Ax - lam[:, None] * x
        Ar = A(r)
        if p is None:
            XX, XAX = G(x, x), G(x, Ax)
            XR, RR, XAR, RAR = G(x, r), G(r, r), G(x, Ar), G(r, Ar)
            S = np.block([[XX, XR], [XR.conj().T, RR]])
            Hm = np.block([[XAX, XAR], [XAR.conj().T, RAR]])
        else:
            XX, XAX, XP, XAP, PP, PAP = (rec[k] for k in ("XX", "XAX", "XP", "XAP", "PP", "PAP"))
            XR = XAX - XX * lam[None, :]                     # X^H (A X - X L)
            RP = XAP - lam[:, None] * XP  # (A X - X L)^H P, A Hermitian
            RR, XAR, RAR, RAP = G(r, r), G(x, Ar), G(r, Ar), G(r, Ap)
            S = np.block([[XX, XR, XP], [XR.conj().T, RR, RP], [XP.conj().T, RP.conj().T, PP]])
            Hm = np.block([[XAX, XAR, XAP], [XAR.conj().T, RAR, RAP], [XAP.conj().T, RAP.conj().T, PAP]])
        S = 0.5 * (S + S.conj().T)
        Hm = 0.5 * (Hm + Hm.conj().T)
        w, C = scipy.linalg.eigh(Hm, S, driver="gvd")
        lam, C = w[:N], C[:, :N]
        Cx, Cr = C[:N], C[N:2 * N]
        Cp = C[2 * N:] if p is not None else None
        pn = Cr.T @ r + (Cp.T @ p if p is not None else 0)
        Apn = Cr.T @ Ar + (Cp.T @ Ap if p is not None else 0)
        xn, Axn = Cx.T @ x + pn, Cx.T @ Ax + Apn
        # recurrence blocks of the next step (NextMatrix_z_blasmp, 1104-1436)
        Cp0 = np.concatenate([np.zeros((N, N), dtype=dt), C[N:]], axis=0)
        Cb = np.concatenate([C, Cp0], axis=1)
        SS, HH = np.conj(Cb).T @ (S @ Cb), np.conj(Cb).T @ (Hm @ Cb)
        rec = {"XX": SS[:N, :N], "XP": SS[:N, N:], "PP": SS[N:, N:],
               "XAX": HH[:N, :N], "XAP": HH[:N, N:], "PAP": HH[N:, N:]}
        x, Ax, p, Ap = xn, Axn, pn, Apn
        history.append(lam.copy())
    return history
