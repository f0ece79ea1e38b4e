/*
 * qm_oracle.c -- CPU FP64 restatement of QUMASUN's H.psi hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; it is the checker that the HIP path is compared against
 * (tests/, __graft_entry__.smoke(), and the cpu_baseline leg of bench.py).
 *
 * Each function names the reference file:line it restates (paths relative to
 * the QUMASUN source tree, source/...).  Arithmetic grouping follows the
 * reference loops so that FP64 results agree to rounding.
 *
 * Layout (source/qumasun_base1.h:451-462, initialize.h:340):
 *   grid index i = ix + Nx*(iy + Ny*iz), x fastest.
 *   band block, ncomp = 2 (SoA complex): band n re at psi + (2n)*Ng,
 *                                         im at psi + (2n+1)*Ng.
 *   band block, ncomp = 1 (real Gamma point): band n at psi + n*Ng.
 *
 * Projector set (per atom a):
 *   nproj[a]            M_a projectors (up to 64; ordered by radial projector k, then
 *                       m = -l_k..l_k, as num_proj_RY in ..._kpoint.h:940-945, 960-966)
 *   vox_off[a]..[a+1]   range of this atom's voxels in the global voxel list
 *   vox_xyz[3v..3v+2]   UNWRAPPED integer grid coordinates of voxel v; the
 *                       grid index is the periodic wrap, the Bloch phase uses
 *                       the unwrapped value (PseudoPotNonlocal_SvF2_kpoint.h:3578-3596)
 *   table[tab_off[a] + w*V_a + v]  projector w of atom a at voxel v
 *                       (column-major V_a x M_a, ..._kpoint.h:2255-2272)
 *   energy[proj_off[a] + w]        E_w (Ha), the KB energy of projector w
 *
 * What pins each function (closed forms, dense matrices, scipy, finite
 * differences) is tested in tests/test_oracle.py; every function has a pin.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int wrap(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

/* ------------------------------------------------------------------------ */
/* FD8 k-point kinetic operator, accumulated (+=) into out.
 * Restates LaplacianKpoint8th_halo_f<KX,KY,KZ> (GridDifferenceKpoint8th.h:17-120)
 * with the periodic halo of PasteHalo_f4 (GridDifference8th.h:18-78) expressed
 * as a modular index, and the re/im coupling of LaplacianKpoint8th_halo
 * (GridDifferenceKpoint8th.h:124-151): the re source is applied with
 * sign1st=+1 writing the first-derivative term into im, the im source with
 * sign1st=-1 writing into re.  coef = -1/2 (qumasun_base1.matrix.h:309).
 * g = k vector in 1/bohr (kpoint index * 2pi/(L*Nk), constructor.h:32-34).
 */
static void fd8_component(int nx, int ny, int nz, double hx, double hy, double hz,
                          double coef, double sign1st, double gx, double gy, double gz,
                          const double *p, double *out, double *out2) {
    const double c0 = -14350.0 / 5040.0 * coef * (1.0 / (hx * hx) + 1.0 / (hy * hy) + 1.0 / (hz * hz))
                      - coef * (gx * gx + gy * gy + gz * gz);
    const double cx[5] = {0.0, 8064.0 * coef / (5040.0 * hx * hx), -1008.0 * coef / (5040.0 * hx * hx),
                          128.0 * coef / (5040.0 * hx * hx), -9.0 * coef / (5040.0 * hx * hx)};
    const double cy[5] = {0.0, 8064.0 * coef / (5040.0 * hy * hy), -1008.0 * coef / (5040.0 * hy * hy),
                          128.0 * coef / (5040.0 * hy * hy), -9.0 * coef / (5040.0 * hy * hy)};
    const double cz[5] = {0.0, 8064.0 * coef / (5040.0 * hz * hz), -1008.0 * coef / (5040.0 * hz * hz),
                          128.0 * coef / (5040.0 * hz * hz), -9.0 * coef / (5040.0 * hz * hz)};
    const double coef2 = sign1st * coef * 2.0;
    const double ax[5] = {0.0, 672.0 * coef2 * gx / (840.0 * hx), -168.0 * coef2 * gx / (840.0 * hx),
                          32.0 * coef2 * gx / (840.0 * hx), -3.0 * coef2 * gx / (840.0 * hx)};
    const double ay[5] = {0.0, 672.0 * coef2 * gy / (840.0 * hy), -168.0 * coef2 * gy / (840.0 * hy),
                          32.0 * coef2 * gy / (840.0 * hy), -3.0 * coef2 * gy / (840.0 * hy)};
    const double az[5] = {0.0, 672.0 * coef2 * gz / (840.0 * hz), -168.0 * coef2 * gz / (840.0 * hz),
                          32.0 * coef2 * gz / (840.0 * hz), -3.0 * coef2 * gz / (840.0 * hz)};
    /* the reference dispatches on |g_d| < 1e-8 (GridDifferenceKpoint8th.h:330-365) */
    const int kx_on = fabs(gx) >= 1.0e-8, ky_on = fabs(gy) >= 1.0e-8, kz_on = fabs(gz) >= 1.0e-8;
    const int any_k = kx_on || ky_on || kz_on;

    for (int iz = 0; iz < nz; ++iz) {
        int zm[5], zp[5];
        for (int j = 1; j <= 4; ++j) { zm[j] = wrap(iz - j, nz); zp[j] = wrap(iz + j, nz); }
        for (int iy = 0; iy < ny; ++iy) {
            int ym[5], yp[5];
            for (int j = 1; j <= 4; ++j) { ym[j] = wrap(iy - j, ny); yp[j] = wrap(iy + j, ny); }
            for (int ix = 0; ix < nx; ++ix) {
                int xm[5], xp[5];
                for (int j = 1; j <= 4; ++j) { xm[j] = wrap(ix - j, nx); xp[j] = wrap(ix + j, nx); }
                const int64_t row = (int64_t)nx * (iy + (int64_t)ny * iz);
#define PX(j) p[xp[j] + row]
#define MX(j) p[xm[j] + row]
#define PY(j) p[ix + (int64_t)nx * (yp[j] + (int64_t)ny * iz)]
#define MY(j) p[ix + (int64_t)nx * (ym[j] + (int64_t)ny * iz)]
#define PZ(j) p[ix + (int64_t)nx * (iy + (int64_t)ny * zp[j])]
#define MZ(j) p[ix + (int64_t)nx * (iy + (int64_t)ny * zm[j])]
                const int64_t io = ix + row;
                double d2 = c0 * p[io];
                d2 += cx[4] * (PX(4) + MX(4)) + cx[3] * (PX(3) + MX(3)) + cx[2] * (PX(2) + MX(2)) + cx[1] * (PX(1) + MX(1));
                d2 += cy[4] * (PY(4) + MY(4)) + cy[3] * (PY(3) + MY(3)) + cy[2] * (PY(2) + MY(2)) + cy[1] * (PY(1) + MY(1));
                d2 += cz[4] * (PZ(4) + MZ(4)) + cz[3] * (PZ(3) + MZ(3)) + cz[2] * (PZ(2) + MZ(2)) + cz[1] * (PZ(1) + MZ(1));
                out[io] += d2;
                if (any_k && out2) {
                    double d1 = 0.0;
                    if (kx_on) d1 += ax[4] * (PX(4) - MX(4)) + ax[3] * (PX(3) - MX(3)) + ax[2] * (PX(2) - MX(2)) + ax[1] * (PX(1) - MX(1));
                    if (ky_on) d1 += ay[4] * (PY(4) - MY(4)) + ay[3] * (PY(3) - MY(3)) + ay[2] * (PY(2) - MY(2)) + ay[1] * (PY(1) - MY(1));
                    if (kz_on) d1 += az[4] * (PZ(4) - MZ(4)) + az[3] * (PZ(3) - MZ(3)) + az[2] * (PZ(2) - MZ(2)) + az[1] * (PZ(1) - MZ(1));
                    out2[io] += d1;
                }
#undef PX
#undef MX
#undef PY
#undef MY
#undef PZ
#undef MZ
            }
        }
    }
}

/* Kinetic part only, accumulated: out += -1/2 (Delta + 2ik.grad - k^2) psi.
 * mKineticMatrixAdd_ddm_kpint (qumasun_base1.matrix.h:298-317). */
void oracle_kinetic_add(int nx, int ny, int nz, double hx, double hy, double hz,
                        int nb, int ncomp, const double *psi, double *out,
                        double gx, double gy, double gz) {
    const int64_t ng = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < nb; ++n) {
        if (ncomp == 2) {
            const double *re = psi + (2 * (int64_t)n) * ng, *im = re + ng;
            double *ore = out + (2 * (int64_t)n) * ng, *oim = ore + ng;
            fd8_component(nx, ny, nz, hx, hy, hz, -0.5, +1.0, gx, gy, gz, re, ore, oim);
            fd8_component(nx, ny, nz, hx, hy, hz, -0.5, -1.0, gx, gy, gz, im, oim, ore);
        } else {
            fd8_component(nx, ny, nz, hx, hy, hz, -0.5, +1.0, 0.0, 0.0, 0.0,
                          psi + (int64_t)n * ng, out + (int64_t)n * ng, NULL);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Nonlocal KB apply, accumulated into hpsi.
 * Restates ProjectionPP_bundle_1 (PseudoPotNonlocal_SvF2_kpoint.h:3910-4046)
 * with mInnerNonlocalPsi_v4_bundle (..._kpoint.h:4463-4576):
 *   cut psi on the atom's voxels; multiply by e^{+ik.r} (4513-4527);
 *   inner = dV * P^T cut (blas_DGEMM_t with alpha=dV, 4531);
 *   c_w = E_w * inner_w (3986-3995);
 *   cutHp = P c (blas_DGEMM_n, 4003); multiply by e^{-ik.r} (4010-4025);
 *   add back onto the grid (AddSubgridByRanges, 4038-4041).
 * Phase k.r = gx*hx*ix + gy*hy*iy + gz*hz*iz on the unwrapped coordinates
 * (mCreateBlochExp, ..._kpoint.h:3578-3596; gx_dx = k*dk*dx, constructor.h:481).
 */
void oracle_nonlocal_add(int nx, int ny, int nz, double hx, double hy, double hz,
                         int nb, int ncomp, const double *psi, double *hpsi,
                         double gx, double gy, double gz,
                         int natoms, const int *nproj, const int64_t *vox_off, const int *vox_xyz,
                         const double *table, const int64_t *tab_off, const double *energy,
                         const int *proj_off) {
    const int64_t ng = (int64_t)nx * ny * nz;
    const double dV = hx * hy * hz;
    const int is_gamma = (gx == 0.0 && gy == 0.0 && gz == 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < nb; ++n) {
        for (int a = 0; a < natoms; ++a) {
            const int64_t v0 = vox_off[a], V = vox_off[a + 1] - vox_off[a];
            const int M = nproj[a];
            if (V == 0 || M == 0) continue;
            const double *P = table + tab_off[a];
            double *cre = (double *)malloc(sizeof(double) * V * 2);
            double *cim = cre + V;
            int64_t *gidx = (int64_t *)malloc(sizeof(int64_t) * V);
            double *cs = (double *)malloc(sizeof(double) * V * 2);
            for (int64_t v = 0; v < V; ++v) {
                const int ix = vox_xyz[3 * (v0 + v)], iy = vox_xyz[3 * (v0 + v) + 1], iz = vox_xyz[3 * (v0 + v) + 2];
                gidx[v] = wrap(ix, nx) + (int64_t)nx * (wrap(iy, ny) + (int64_t)ny * wrap(iz, nz));
                const double ph = gx * hx * (double)ix + gy * hy * (double)iy + gz * hz * (double)iz;
                cs[2 * v] = cos(ph);
                cs[2 * v + 1] = sin(ph);
            }
            const double *pre, *pim = NULL;
            if (ncomp == 2) { pre = psi + 2 * (int64_t)n * ng; pim = pre + ng; }
            else { pre = psi + (int64_t)n * ng; }
            for (int64_t v = 0; v < V; ++v) {
                const double r = pre[gidx[v]], i = pim ? pim[gidx[v]] : 0.0;
                if (!is_gamma) {
                    cre[v] = cs[2 * v] * r - cs[2 * v + 1] * i;
                    cim[v] = cs[2 * v] * i + cs[2 * v + 1] * r;
                } else {
                    cre[v] = r;
                    cim[v] = i;
                }
            }
            double cr[64], ci[64];  /* M <= 64 (the library takes at most 32) */
            if (M > 64) abort();
            for (int w = 0; w < M; ++w) {
                double sr = 0.0, si = 0.0;
                for (int64_t v = 0; v < V; ++v) {
                    sr += P[w * V + v] * cre[v];
                    si += P[w * V + v] * cim[v];
                }
                cr[w] = dV * sr * energy[proj_off[a] + w];
                ci[w] = dV * si * energy[proj_off[a] + w];
            }
            double *hre, *him = NULL;
            if (ncomp == 2) { hre = hpsi + 2 * (int64_t)n * ng; him = hre + ng; }
            else { hre = hpsi + (int64_t)n * ng; }
            for (int64_t v = 0; v < V; ++v) {
                double sr = 0.0, si = 0.0;
                for (int w = 0; w < M; ++w) {
                    sr += P[w * V + v] * cr[w];
                    si += P[w * V + v] * ci[w];
                }
                if (!is_gamma) {
                    const double tr = cs[2 * v] * sr + cs[2 * v + 1] * si;
                    const double ti = cs[2 * v] * si - cs[2 * v + 1] * sr;
                    sr = tr;
                    si = ti;
                }
                hre[gidx[v]] += sr;
                if (him) him[gidx[v]] += si;
            }
            free(cre);
            free(gidx);
            free(cs);
        }
    }
}

/* Full H.psi of the FD8 assembly (qumasun_base1.matrix.h:262-291):
 * Hpsi = 0; Hpsi = V psi (re and im); Hpsi += Vnl psi; Hpsi += K psi. */
void oracle_hpsi(int nx, int ny, int nz, double hx, double hy, double hz,
                 const double *vloc, int nb, int ncomp, const double *psi, double *hpsi,
                 double gx, double gy, double gz,
                 int natoms, const int *nproj, const int64_t *vox_off, const int *vox_xyz,
                 const double *table, const int64_t *tab_off, const double *energy,
                 const int *proj_off) {
    const int64_t ng = (int64_t)nx * ny * nz;
    const int64_t ncol = (int64_t)nb * ncomp;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncol; ++c) {
        const double *s = psi + c * ng;
        double *d = hpsi + c * ng;
        for (int64_t i = 0; i < ng; ++i) d[i] = vloc[i] * s[i];
    }
    if (natoms > 0)
        oracle_nonlocal_add(nx, ny, nz, hx, hy, hz, nb, ncomp, psi, hpsi, gx, gy, gz,
                            natoms, nproj, vox_off, vox_xyz, table, tab_off, energy, proj_off);
    oracle_kinetic_add(nx, ny, nz, hx, hy, hz, nb, ncomp, psi, hpsi, gx, gy, gz);
}

/* ------------------------------------------------------------------------ */
/* Real spherical harmonics and the synthetic KB projector.
 * Y_lm = c_lm P_l^{|m|}(z) T_m(x,y) on the unit vector (RealSphericalHarmonics.h:148-162,
 * Legendre.h:97-144 with s = 1, Trigonometric.h:78-112); no Condon-Shortley sign.
 * c_lm = sqrt((2l+1)(l-|m|)!/(2 pi (l+|m|)!)) for m != 0, sqrt((2l+1)/(4 pi)) for m = 0.
 * At r = 0: Y00 = 1/sqrt(4 pi), others 0 (SetBlockYlm.h:534-539).
 * Ordering inside l: m = -l..l (SetBlockYlm.h:666-740).
 */
static double fact(int n) { double f = 1.0; for (int i = 2; i <= n; ++i) f *= i; return f; }

double oracle_ylm(int l, int m, double x, double y, double z) {
    const double rr = x * x + y * y + z * z;
    if (rr < 1.0e-14) return (l == 0) ? 1.0 / sqrt(4.0 * M_PI) : 0.0;
    const double r = sqrt(rr);
    const double ux = x / r, uy = y / r, uz = z / r;
    const int am = m < 0 ? -m : m;
    double c;
    if (m != 0) c = sqrt((double)(2 * l + 1) * fact(l - am) / (2.0 * fact(l + am) * M_PI));
    else c = sqrt((double)(2 * l + 1) / (4.0 * M_PI));
    double P = 1.0; /* Plm(z, s=1) */
    if (l == 1 && am == 0) P = uz;
    else if (l == 1 && am == 1) P = 1.0;
    else if (l == 2 && am == 0) P = 0.5 * (3.0 * uz * uz - 1.0);
    else if (l == 2 && am == 1) P = 3.0 * uz;
    else if (l == 2 && am == 2) P = 3.0;
    else if (l == 3 && am == 0) P = 0.5 * (5.0 * uz * uz * uz - 3.0 * uz);
    else if (l == 3 && am == 1) P = 1.5 * (5.0 * uz * uz - 1.0);
    else if (l == 3 && am == 2) P = 15.0 * uz;
    else if (l == 3 && am == 3) P = 15.0;
    /* T_m = Re (ux + i uy)^m for m > 0, Im (ux + i uy)^|m| for m < 0 */
    double T = 1.0;
    if (m == 1) T = ux;
    else if (m == -1) T = uy;
    else if (m == 2) T = ux * ux - uy * uy;
    else if (m == -2) T = 2.0 * ux * uy;
    else if (m == 3) T = ux * ux * ux - 3.0 * ux * uy * uy;
    else if (m == -3) T = 3.0 * ux * ux * uy - uy * uy * uy;
    return c * P * T;
}

/* Radial projector irad of angular momentum l:
 *   irad 0: p_lm(r) = r^l exp(-r^2 / (2 sigma^2)) Y_lm(r_hat)
 *   irad 1: p_lm(r) = r^l (r^2 / sigma^2) exp(-r^2 / (2 sigma^2)) Y_lm(r_hat)
 * for r < rc, else 0. */
double oracle_kb_projector_radial(int l, int m, int irad, double sigma, double rc, double x, double y, double z) {
    const double rr = x * x + y * y + z * z;
    if (rr >= rc * rc) return 0.0;
    const double r = sqrt(rr);
    double rl = 1.0;
    for (int i = 0; i < l; ++i) rl *= r;
    const double u = irad == 1 ? rr / (sigma * sigma) : 1.0;
    return rl * u * exp(-rr / (2.0 * sigma * sigma)) * oracle_ylm(l, m, x, y, z);
}

double oracle_kb_projector(int l, int m, double sigma, double rc, double x, double y, double z) {
    return oracle_kb_projector_radial(l, m, 0, sigma, rc, x, y, z);
}


/* Projector table of one atom set: table[tab_off[a] + w*V_a + v] = p_{l(w) m(w)}(r_v - R_a)
 * with w enumerating l = 0..lmax[a], then the nradial radial projectors of that l, then
 * m = -l..l (the reference's order: radial projector k, then m, ..._kpoint.h:960-966,
 * SetBlockYlm.h:666-740) and r_v = (ix*hx, iy*hy, iz*hz) on the unwrapped voxel coordinates. */
void oracle_projector_table(int natoms, const double *pos, const int *lmax, const double *rc, int nradial,
                            double sigma, double hx, double hy, double hz,
                            const int64_t *vox_off, const int *vox_xyz, const int64_t *tab_off,
                            double *table) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int a = 0; a < natoms; ++a) {
        const int64_t v0 = vox_off[a], V = vox_off[a + 1] - vox_off[a];
        int w = 0;
        for (int l = 0; l <= lmax[a]; ++l)
            for (int k = 0; k < nradial; ++k)
                for (int m = -l; m <= l; ++m, ++w) {
                    for (int64_t v = 0; v < V; ++v) {
                        const double x = hx * vox_xyz[3 * (v0 + v)] - pos[3 * a];
                        const double y = hy * vox_xyz[3 * (v0 + v) + 1] - pos[3 * a + 1];
                        const double z = hz * vox_xyz[3 * (v0 + v) + 2] - pos[3 * a + 2];
                        table[tab_off[a] + w * V + v] = oracle_kb_projector_radial(l, m, k, sigma, rc[a], x, y, z);
                    }
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Adjoint w.r.t. V_loc: band density rho(r) += sum_n f_n (re^2 + im^2).
 * SetDensityOne (qumasun_SetDensityOne.h:26-39); the gradient of
 * E_band = sum_n f_n <psi_n|H|psi_n> dV with respect to V_loc(r) is rho(r) dV. */
void oracle_band_density(int64_t ng, int nb, int ncomp, const double *psi, const double *occ, double *rho) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ng; ++i) {
        double s = 0.0;
        for (int c = 0; c < nb * ncomp; ++c) {
            const double v = psi[i + ng * (int64_t)c];
            s += (v * v) * occ[c / ncomp];
        }
        rho[i] += s;
    }
}

/* Nonlocal energy and forces.  Restates ForceNonlocal_bundle with
 * FORCE_DIFF_PROJ (PseudoPotNonlocal_SvF2_kpoint.h:3629-3731) and
 * mInnerNonlocalPsi_diff_bundle (4775-4880):
 *   inner  = dV sum_v p_w(v) e^{ik.r} psi(v),   diff_d = dV sum_v d_d p_w(v) e^{ik.r} psi(v)
 *   E_nl  += f_n E_w |inner|^2
 *   F_a,d += 2 f_n E_w Re(conj(inner) diff_d)          (= -dE_nl/dR_a,d)
 * dtable holds d p / d r_d (gradient with respect to the grid point) in the
 * same layout as table, for d = x, y, z consecutively: dtable[d*tab_total + ...]. */
double oracle_nonlocal_forces(int nx, int ny, int nz, double hx, double hy, double hz,
                              int nb, int ncomp, const double *psi, const double *occ,
                              double gx, double gy, double gz,
                              int natoms, const int *nproj, const int64_t *vox_off, const int *vox_xyz,
                              const double *table, const double *dtable, int64_t tab_total,
                              const int64_t *tab_off, const double *energy, const int *proj_off,
                              double *forces) {
    const int64_t ng = (int64_t)nx * ny * nz;
    const double dV = hx * hy * hz;
    double e_nl = 0.0;
    for (int a = 0; a < natoms; ++a) {
        forces[3 * a] = forces[3 * a + 1] = forces[3 * a + 2] = 0.0;
    }
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : e_nl)
    for (int a = 0; a < natoms; ++a) {
        const int64_t v0 = vox_off[a], V = vox_off[a + 1] - vox_off[a];
        const int M = nproj[a];
        double fa[3] = {0.0, 0.0, 0.0};
        for (int n = 0; n < nb; ++n) {
            const double *pre, *pim = NULL;
            if (ncomp == 2) { pre = psi + 2 * (int64_t)n * ng; pim = pre + ng; }
            else { pre = psi + (int64_t)n * ng; }
            for (int w = 0; w < M; ++w) {
                double s[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
                for (int64_t v = 0; v < V; ++v) {
                    const int ix = vox_xyz[3 * (v0 + v)], iy = vox_xyz[3 * (v0 + v) + 1], iz = vox_xyz[3 * (v0 + v) + 2];
                    const int64_t g = wrap(ix, nx) + (int64_t)nx * (wrap(iy, ny) + (int64_t)ny * wrap(iz, nz));
                    const double ph = gx * hx * (double)ix + gy * hy * (double)iy + gz * hz * (double)iz;
                    const double c = cos(ph), sn = sin(ph);
                    const double r = pre[g], i = pim ? pim[g] : 0.0;
                    const double br = c * r - sn * i, bi = c * i + sn * r;
                    const int64_t t = tab_off[a] + w * V + v;
                    const double pv[4] = {table[t], dtable[t], dtable[tab_total + t], dtable[2 * tab_total + t]};
                    for (int q = 0; q < 4; ++q) { s[q][0] += pv[q] * br; s[q][1] += pv[q] * bi; }
                }
                for (int q = 0; q < 4; ++q) { s[q][0] *= dV; s[q][1] *= dV; }
                const double eo = energy[proj_off[a] + w] * occ[n];
                e_nl += (s[0][0] * s[0][0] + s[0][1] * s[0][1]) * eo;
                for (int d = 0; d < 3; ++d)
                    fa[d] += 2.0 * (s[0][0] * s[d + 1][0] + s[0][1] * s[d + 1][1]) * eo;
            }
        }
        forces[3 * a] = fa[0];
        forces[3 * a + 1] = fa[1];
        forces[3 * a + 2] = fa[2];
    }
    return e_nl;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
