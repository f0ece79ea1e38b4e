// UPF pseudopotential loader and real-space projector tables (host code).
//
// Loading follows LoadPseudoPotential_UPF_ONCV (upf_oncv_loader.h:355-530) for
// the nonlocal part: number_of_proj and z_valence from PP_HEADER, the radial
// mesh PP_R (taken as uniform: r0 = r[0], dr = r[1] - r[0]), each PP_BETA.k
// with its angular_momentum and cutoff_radius, and the diagonal of PP_DIJ
// halved (Ry -> Ha) as the KB energy of projector k.  The radial function is
// resampled in two stages, as the reference does:
//  1. beta_k(r) (stored times r) is divided by r, with the r = 0 value taken
//     from the next point for l = 0 and 0 otherwise, and interpolated at the
//     points of a logarithmic grid r_i = exp(xi_min + i dxi), i < 1000,
//     xi_min = -9, dxi = (log r_last - xi_min) / 1000, by UniformGrid::GetValue
//     (UniformGrid.h:40-87): a quintic between mesh points with value, first
//     and second derivative at both ends, the derivatives from five-point
//     central differences, the mesh mirrored through r = 0 (even for even l,
//     odd for odd l) and zero past its end;
//  2. at a grid point, f_k(r) is RadialGrid2::GetValueByXi at xi = log r
//     (RadialGrid2.h:81-157: the same quintic in the interior, a cubic from
//     central differences at the second and third-to-last cells, linear in the
//     first cell, zero from the second-to-last on), f_k(r_0) below r_0, and 0
//     beyond the largest cutoff radius of the projectors of that l
//     (PseudoPotNonlocal_SvF2_kpoint.h:1003-1024).
//
// Tables: for atoms at given positions, every grid voxel with |r_v - R| below
// the largest cutoff radius (unwrapped integer coordinates, ordered z, y, x),
// projector k, m = -l..l (the reference's order, num_proj_RY,
// PseudoPotNonlocal_SvF2_kpoint.h:940-966) valued
//   p_{k l m}(r) = f_k(|r|) Y_lm(r / |r|),
// with Y_lm the reference's real harmonics (RealSphericalHarmonics.h:148-162,
// no Condon-Shortley sign) written as solid harmonics S_lm = r^l Y_lm, and the
// gradient tables d p / d r from the derivative of the same interpolants.  The
// reference then filters the tables in Fourier space on a refined subgrid
// (mInitializeNonlocalProjector / MakeProjectorInRealspace,
// ..._kpoint.h:911-1178, 1912-2208); that step is not restated here: these are
// the unfiltered point values.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "qm_internal.h"

struct qm_upf {
    int nradial = 0;
    double z_valence = 0.0;
    double r0 = 0.0, dr = 0.0;  // uniform UPF mesh
    int nmesh = 0;
    double xi_min = -9.0, dxi = 0.0;  // logarithmic resampling grid
    std::vector<int> l;
    std::vector<double> cutoff, energy;
    std::vector<std::vector<double>> f;    // beta_k / r on the UPF mesh
    std::vector<std::vector<double>> flog; // f_k on the logarithmic grid
    double cutoff_l[4] = {0.0, 0.0, 0.0, 0.0};
    double rc_max = 0.0;
};

namespace qm {
namespace {

// The text between "<tag" and the next '>' (attributes), or "" if absent.
bool find_tag(const std::string &s, const std::string &tag, size_t *attr_begin, size_t *attr_end, size_t *body) {
    size_t pos = 0;
    const std::string open = "<" + tag;
    while ((pos = s.find(open, pos)) != std::string::npos) {
        const size_t after = pos + open.size();
        if (after < s.size() && (std::isspace((unsigned char)s[after]) || s[after] == '>' || s[after] == '/')) {
            const size_t gt = s.find('>', after);
            if (gt == std::string::npos) return false;
            *attr_begin = after;
            *attr_end = gt;
            *body = gt + 1;
            return true;
        }
        pos = after;
    }
    return false;
}

bool attribute(const std::string &s, size_t b, size_t e, const std::string &name, std::string *val) {
    size_t pos = b;
    while (true) {
        pos = s.find(name, pos);
        if (pos == std::string::npos || pos >= e) return false;
        size_t q = pos + name.size();
        while (q < e && std::isspace((unsigned char)s[q])) ++q;
        const bool word_start = pos == b || std::isspace((unsigned char)s[pos - 1]);
        if (word_start && q < e && s[q] == '=') {
            q = s.find('"', q);
            if (q == std::string::npos || q >= e) return false;
            const size_t q2 = s.find('"', q + 1);
            if (q2 == std::string::npos || q2 > e) return false;
            *val = s.substr(q + 1, q2 - q - 1);
            return true;
        }
        pos += name.size();
    }
}

// count numbers from the body of tag (up to its closing tag); missing values read as 0
std::vector<double> numbers(const std::string &s, size_t body, const std::string &tag, int count) {
    const size_t end = s.find("</" + tag, body);
    const std::string text = s.substr(body, end == std::string::npos ? std::string::npos : end - body);
    std::vector<double> v;
    v.reserve(count);
    const char *p = text.c_str();
    for (int i = 0; i < count; ++i) {
        char *q = nullptr;
        const double d = std::strtod(p, &q);
        if (q == p) {
            v.push_back(0.0);
            continue;
        }
        v.push_back(d);
        p = q;
    }
    return v;
}

constexpr int kLogGrid = 1000;

// quintic on [0, 1] with value, first and second derivative d0 at 0 and d1 at
// 1 (derivatives per cell): value and d/dx at x
void quintic(double x, double f0, double d1f0, double d2f0, double f1, double d1f1, double d2f1, double *v,
             double *dv) {
    const double e0 = f1 - (f0 + d1f0 + d2f0 / 2.0);
    const double e1 = d1f1 - (d1f0 + d2f0);
    const double e2 = d2f1 - d2f0;
    const double c3 = 10.0 * e0 - 4.0 * e1 + 0.5 * e2;
    const double c4 = -15.0 * e0 + 7.0 * e1 - e2;
    const double c5 = e0 - c3 - c4;
    *v = f0 + x * (d1f0 + x * (d2f0 / 2.0 + x * (c3 + x * (c4 + x * c5))));
    *dv = d1f0 + x * (d2f0 + x * (3.0 * c3 + x * (4.0 * c4 + x * 5.0 * c5)));
}

// five-point central first and second derivatives at the two ends of cell i
void cell_derivs(double fm2, double fm1, double f0, double f1, double f2, double f3, double *d1f0, double *d1f1,
                 double *d2f0, double *d2f1) {
    *d1f0 = 2.0 / 3.0 * (f1 - fm1) - 1.0 / 12.0 * (f2 - fm2);
    *d1f1 = 2.0 / 3.0 * (f2 - f0) - 1.0 / 12.0 * (f3 - fm1);
    *d2f0 = 4.0 / 3.0 * (f1 + fm1) - 1.0 / 12.0 * (f2 + fm2) - 5.0 / 2.0 * f0;
    *d2f1 = 4.0 / 3.0 * (f2 + f0) - 1.0 / 12.0 * (f3 + fm1) - 5.0 / 2.0 * f1;
}

// UniformGrid::GetValue on the mesh r0 + i dr (mirror parity `par` through the
// first point, zero past the end): the value at r
double mesh_value(const std::vector<double> &vals, double r0, double dr, int par, double r) {
    const int n = (int)vals.size();
    const double rel = (r - r0) / dr;
    const double fl = std::floor(rel);
    const int i = (int)fl;
    const double x = rel - fl;
    auto get = [&](int j) {
        if (j < 0) return vals[-j] * (double)par;
        if (j >= n) return 0.0;
        return vals[j];
    };
    double d1f0, d1f1, d2f0, d2f1, v, dv;
    cell_derivs(get(i - 2), get(i - 1), get(i), get(i + 1), get(i + 2), get(i + 3), &d1f0, &d1f1, &d2f0, &d2f1);
    quintic(x, get(i), d1f0, d2f0, get(i + 1), d1f1, d2f1, &v, &dv);
    return v;
}

// RadialGrid2::GetValueByXi on the logarithmic grid: value and d/dxi at xi
void xi_value(const std::vector<double> &v, double xi_min, double dxi, double xi, double *f, double *dfdxi) {
    const int n = (int)v.size();
    const double rel = (xi - xi_min) / dxi;
    const double fl = std::floor(rel);
    const int i = (int)fl;
    const double x = rel - fl;
    *f = 0.0;
    *dfdxi = 0.0;
    if (i >= n - 2 || i < 0) return;
    double val, dval;
    if (i == 0) {
        val = v[0] * (1.0 - x) + v[1] * x;
        dval = v[1] - v[0];
    } else if (i > 1 && i < n - 3) {
        double d1f0, d1f1, d2f0, d2f1;
        cell_derivs(v[i - 2], v[i - 1], v[i], v[i + 1], v[i + 2], v[i + 3], &d1f0, &d1f1, &d2f0, &d2f1);
        quintic(x, v[i], d1f0, d2f0, v[i + 1], d1f1, d2f1, &val, &dval);
    } else {  // i == 1 or i == n - 3: cubic with central-difference slopes
        const double v0 = v[i], v1 = v[i + 1];
        const double s1 = (v[i + 2] - v0) / 2.0, s0 = (v1 - v[i - 1]) / 2.0;
        const double a = s1 + s0 - 2.0 * (v1 - v0), b = -s1 - 2.0 * s0 + 3.0 * (v1 - v0);
        val = ((a * x + b) * x + s0) * x + v0;
        dval = (3.0 * a * x + 2.0 * b) * x + s0;
    }
    *f = val;
    *dfdxi = dval / dxi;
}

// the radial function f_k the tables use, and df/dr
void radial_value(const qm_upf *u, int k, double r, double *f, double *df) {
    const double rmin = std::exp(u->xi_min);
    const double rc = u->cutoff_l[u->l[k]];
    if (r < rmin) {
        *f = u->flog[k][0];
        *df = 0.0;
    } else if (r > rc) {
        *f = 0.0;
        *df = 0.0;
    } else {
        double dxi;
        xi_value(u->flog[k], u->xi_min, u->dxi, std::log(r), f, &dxi);
        *df = dxi / r;
    }
}

// Real solid harmonics S_lm = r^l Y_lm (m = -l..l, l <= 3) and their gradients
void solid_harmonic(int l, int m, double x, double y, double z, double *s, double g[3]) {
    const double pi = M_PI;
    g[0] = g[1] = g[2] = 0.0;
    if (l == 0) {
        *s = 1.0 / std::sqrt(4.0 * pi);
        return;
    }
    if (l == 1) {
        const double c = std::sqrt(3.0 / (4.0 * pi));
        if (m == -1) { *s = c * y; g[1] = c; }
        else if (m == 0) { *s = c * z; g[2] = c; }
        else { *s = c * x; g[0] = c; }
        return;
    }
    if (l == 2) {
        const double ca = std::sqrt(15.0 / (4.0 * pi)), cb = std::sqrt(5.0 / (16.0 * pi)),
                     cc = std::sqrt(15.0 / (16.0 * pi));
        switch (m) {
            case -2: *s = ca * x * y; g[0] = ca * y; g[1] = ca * x; break;
            case -1: *s = ca * y * z; g[1] = ca * z; g[2] = ca * y; break;
            case 0: *s = cb * (2 * z * z - x * x - y * y); g[0] = -2 * cb * x; g[1] = -2 * cb * y; g[2] = 4 * cb * z; break;
            case 1: *s = ca * x * z; g[0] = ca * z; g[2] = ca * x; break;
            default: *s = cc * (x * x - y * y); g[0] = 2 * cc * x; g[1] = -2 * cc * y; break;
        }
        return;
    }
    const double a = 0.25 * std::sqrt(35.0 / (2.0 * pi)), b = 0.5 * std::sqrt(105.0 / pi),
                 c = 0.25 * std::sqrt(21.0 / (2.0 * pi)), d = 0.25 * std::sqrt(7.0 / pi),
                 e = 0.25 * std::sqrt(105.0 / pi);
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    switch (m) {
        case -3: *s = a * (3 * x2 * y - y2 * y); g[0] = 6 * a * x * y; g[1] = a * (3 * x2 - 3 * y2); break;
        case -2: *s = b * x * y * z; g[0] = b * y * z; g[1] = b * x * z; g[2] = b * x * y; break;
        case -1: *s = c * y * (4 * z2 - x2 - y2); g[0] = -2 * c * x * y; g[1] = c * (4 * z2 - x2 - 3 * y2); g[2] = 8 * c * y * z; break;
        case 0: *s = d * z * (2 * z2 - 3 * x2 - 3 * y2); g[0] = -6 * d * x * z; g[1] = -6 * d * y * z; g[2] = d * (6 * z2 - 3 * x2 - 3 * y2); break;
        case 1: *s = c * x * (4 * z2 - x2 - y2); g[0] = c * (4 * z2 - 3 * x2 - y2); g[1] = -2 * c * x * y; g[2] = 8 * c * x * z; break;
        case 2: *s = e * z * (x2 - y2); g[0] = 2 * e * x * z; g[1] = -2 * e * y * z; g[2] = e * (x2 - y2); break;
        default: *s = a * (x2 * x - 3 * x * y2); g[0] = a * (3 * x2 - 3 * y2); g[1] = -6 * a * x * y; break;
    }
}

int load(qm_upf *u, const char *path) {
    std::ifstream in(path, std::ios::binary);
    QM_CHECK(in.good(), QM_ERR_ARG, std::string("qm_upf_load: cannot open ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string s = ss.str();
    size_t ab, ae, body;
    std::string v;
    QM_CHECK(find_tag(s, "PP_HEADER", &ab, &ae, &body), QM_ERR_ARG, "qm_upf_load: no PP_HEADER (UPF v2 expected)");
    QM_CHECK(attribute(s, ab, ae, "number_of_proj", &v), QM_ERR_ARG, "qm_upf_load: no number_of_proj");
    u->nradial = std::stoi(v);
    QM_CHECK(u->nradial >= 0 && u->nradial <= 32, QM_ERR_ARG, "qm_upf_load: bad number_of_proj");
    if (attribute(s, ab, ae, "z_valence", &v)) u->z_valence = std::stod(v);
    QM_CHECK(find_tag(s, "PP_R", &ab, &ae, &body), QM_ERR_ARG, "qm_upf_load: no PP_R");
    QM_CHECK(attribute(s, ab, ae, "size", &v), QM_ERR_ARG, "qm_upf_load: PP_R without size");
    u->nmesh = std::stoi(v);
    QM_CHECK(u->nmesh >= 8, QM_ERR_ARG, "qm_upf_load: radial mesh too short");
    const std::vector<double> r = numbers(s, body, "PP_R", u->nmesh);
    u->r0 = r[0];
    u->dr = r[1] - r[0];
    QM_CHECK(u->dr > 0.0, QM_ERR_ARG, "qm_upf_load: radial mesh not increasing");
    for (int i = 2; i < u->nmesh; ++i)
        QM_CHECK(std::fabs(r[i] - r[0] - i * u->dr) <= 1e-6 * (1.0 + std::fabs(r[i])), QM_ERR_ARG,
                 "qm_upf_load: the radial mesh is not uniform (ONCV meshes are)");
    int nproj_lm = 0;
    for (int k = 0; k < u->nradial; ++k) {
        const std::string tag = "PP_BETA." + std::to_string(k + 1);
        QM_CHECK(find_tag(s, tag, &ab, &ae, &body), QM_ERR_ARG, "qm_upf_load: no " + tag);
        QM_CHECK(attribute(s, ab, ae, "angular_momentum", &v), QM_ERR_ARG, "qm_upf_load: " + tag + " without l");
        const int l = std::stoi(v);
        QM_CHECK(l >= 0 && l <= 3, QM_ERR_ARG, "qm_upf_load: angular momentum above 3");
        double rc = 0.0;
        if (attribute(s, ab, ae, "cutoff_radius", &v)) rc = std::stod(v);
        std::vector<double> b = numbers(s, body, tag, u->nmesh);
        for (int i = 1; i < u->nmesh; ++i) b[i] /= r[i];
        b[0] = l == 0 ? b[1] : 0.0;
        u->l.push_back(l);
        u->cutoff.push_back(rc);
        u->f.push_back(b);
        u->rc_max = std::max(u->rc_max, rc);
        u->cutoff_l[l] = std::max(u->cutoff_l[l], rc);
        nproj_lm += 2 * l + 1;
    }
    QM_CHECK(nproj_lm <= kMaxProj, QM_ERR_ARG, "qm_upf_load: more than 32 projectors per atom");
    const double r_last = r[u->nmesh - 1];
    QM_CHECK(r_last > std::exp(u->xi_min), QM_ERR_ARG, "qm_upf_load: radial mesh too short");
    u->dxi = (std::log(r_last) - u->xi_min) / (double)kLogGrid;
    for (int k = 0; k < u->nradial; ++k) {
        const int par = (u->l[k] & 1) ? -1 : 1;
        std::vector<double> g(kLogGrid);
        for (int i = 0; i < kLogGrid; ++i)
            g[i] = mesh_value(u->f[k], u->r0, u->dr, par, std::exp((double)i * u->dxi + u->xi_min));
        u->flog.push_back(std::move(g));
    }
    if (u->nradial > 0) {
        QM_CHECK(find_tag(s, "PP_DIJ", &ab, &ae, &body), QM_ERR_ARG, "qm_upf_load: no PP_DIJ");
        const std::vector<double> d = numbers(s, body, "PP_DIJ", u->nradial * u->nradial);
        for (int k = 0; k < u->nradial; ++k) u->energy.push_back(d[k + u->nradial * k] / 2.0);  // Ry -> Ha
    }
    return QM_OK;
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

int qm_upf_load(qm_upf **out, const char *path) {
    QM_CHECK(out && path, QM_ERR_ARG, "qm_upf_load: NULL argument");
    *out = nullptr;
    qm_upf *u = new qm_upf();
    int st = QM_OK;
    try {
        st = load(u, path);
    } catch (const std::exception &e) {
        set_error(std::string("qm_upf_load: malformed number: ") + e.what());
        st = QM_ERR_ARG;
    }
    if (st != QM_OK) {
        delete u;
        return st;
    }
    *out = u;
    return QM_OK;
}

void qm_upf_free(qm_upf *u) { delete u; }

int qm_upf_info(const qm_upf *u, int *nradial, int *nproj, double *z_valence, double *cutoff) {
    QM_CHECK(u, QM_ERR_ARG, "qm_upf_info: NULL pseudopotential");
    int n = 0;
    for (int l : u->l) n += 2 * l + 1;
    if (nradial) *nradial = u->nradial;
    if (nproj) *nproj = n;
    if (z_valence) *z_valence = u->z_valence;
    if (cutoff) *cutoff = u->rc_max;
    return QM_OK;
}

int qm_upf_radial(const qm_upf *u, int k, int *l, double *cutoff, double *energy) {
    QM_CHECK(u && k >= 0 && k < u->nradial, QM_ERR_ARG, "qm_upf_radial: bad projector index");
    if (l) *l = u->l[k];
    if (cutoff) *cutoff = u->cutoff[k];
    if (energy) *energy = u->energy[k];
    return QM_OK;
}

int qm_upf_eval(const qm_upf *u, int k, int n, const double *r, double *f, double *df) {
    QM_CHECK(u && k >= 0 && k < u->nradial && n >= 0 && r && f, QM_ERR_ARG, "qm_upf_eval: bad argument");
    for (int i = 0; i < n; ++i) {
        double v, d;
        radial_value(u, k, r[i], &v, &d);
        f[i] = v;
        if (df) df[i] = d;
    }
    return QM_OK;
}

int qm_upf_tables(const qm_upf *u, int natoms, const double *positions, int nx, int ny, int nz, double hx, double hy,
                  double hz, int64_t *counts, int *nproj, int64_t *vox_off, int *vox_xyz, double *table,
                  double *dtable, int64_t *tab_off, double *energy, int *proj_off) {
    QM_CHECK(u && counts && (natoms == 0 || positions), QM_ERR_ARG, "qm_upf_tables: NULL argument");
    QM_CHECK(natoms >= 0 && nx > 0 && ny > 0 && nz > 0 && hx > 0 && hy > 0 && hz > 0, QM_ERR_ARG,
             "qm_upf_tables: bad grid or atom count");
    int M = 0;
    for (int l : u->l) M += 2 * l + 1;
    const double rc = u->rc_max;
    // voxels of each atom's sphere, ordered z, y, x
    std::vector<std::vector<int>> vox(natoms);
    for (int a = 0; a < natoms; ++a) {
        const double *R = positions + 3 * a;
        const int x0 = (int)std::ceil((R[0] - rc) / hx), x1 = (int)std::floor((R[0] + rc) / hx);
        const int y0 = (int)std::ceil((R[1] - rc) / hy), y1 = (int)std::floor((R[1] + rc) / hy);
        const int z0 = (int)std::ceil((R[2] - rc) / hz), z1 = (int)std::floor((R[2] + rc) / hz);
        for (int iz = z0; iz <= z1; ++iz)
            for (int iy = y0; iy <= y1; ++iy)
                for (int ix = x0; ix <= x1; ++ix) {
                    const double dx = ix * hx - R[0], dy = iy * hy - R[1], dz = iz * hz - R[2];
                    if (dx * dx + dy * dy + dz * dz < rc * rc) {
                        vox[a].push_back(ix);
                        vox[a].push_back(iy);
                        vox[a].push_back(iz);
                    }
                }
    }
    int64_t vtot = 0;
    for (int a = 0; a < natoms; ++a) vtot += (int64_t)vox[a].size() / 3;
    const int64_t ttot = vtot * M, ptot = (int64_t)natoms * M;
    const bool fill = nproj && vox_off && vox_xyz && table && tab_off && energy && proj_off;
    counts[0] = vtot;
    counts[1] = ttot;
    counts[2] = ptot;
    if (!fill) return QM_OK;  // sizes only
    int64_t v0 = 0;
    vox_off[0] = tab_off[0] = 0;
    proj_off[0] = 0;
    for (int a = 0; a < natoms; ++a) {
        const double *R = positions + 3 * a;
        const int64_t V = (int64_t)vox[a].size() / 3;
        nproj[a] = M;
        vox_off[a + 1] = vox_off[a] + V;
        tab_off[a + 1] = tab_off[a] + V * M;
        proj_off[a + 1] = proj_off[a] + M;
        std::memcpy(vox_xyz + 3 * v0, vox[a].data(), sizeof(int) * vox[a].size());
        int w = 0;
        for (int k = 0; k < u->nradial; ++k) {
            const int l = u->l[k];
            for (int m = -l; m <= l; ++m, ++w) {
                energy[proj_off[a] + w] = u->energy[k];
                double *t = table + tab_off[a] + (int64_t)w * V;
                for (int64_t v = 0; v < V; ++v) {
                    const double x = vox[a][3 * v] * hx - R[0], y = vox[a][3 * v + 1] * hy - R[1],
                                 z = vox[a][3 * v + 2] * hz - R[2];
                    const double rr = std::sqrt(x * x + y * y + z * z);
                    double f, df, sv, gs[3];
                    radial_value(u, k, rr, &f, &df);
                    solid_harmonic(l, m, x, y, z, &sv, gs);
                    double p = 0.0, dp[3] = {0.0, 0.0, 0.0};
                    if (rr > 1e-12) {
                        // p = f(r) S / r^l;  grad p = f' (r/|r|) S / r^l + f (grad S / r^l - l S r / r^(l+2))
                        const double il = 1.0 / std::pow(rr, l);
                        const double xyz[3] = {x, y, z};
                        p = f * sv * il;
                        for (int c = 0; c < 3; ++c)
                            dp[c] = df * xyz[c] / rr * sv * il + f * (gs[c] * il - l * sv * il * xyz[c] / (rr * rr));
                    } else if (l == 0) {
                        p = f * sv;
                    }
                    t[v] = p;
                    if (dtable)
                        for (int c = 0; c < 3; ++c) dtable[c * ttot + tab_off[a] + (int64_t)w * V + v] = dp[c];
                }
            }
        }
        v0 += V;
    }
    return QM_OK;
}

}  // extern "C"
