"""Synthetic UPF v2 files for the loader tests: Gaussian projectors
beta_k(r) = r^(l+1) exp(-r^2 / 2 s^2) (stored times r, as UPF does) on the
uniform 0.01-bohr mesh ONCV files use."""
import numpy as np

SIGMA = 0.6
MESH_DR = 0.01
MESH_N = 1401


def gaussian_beta(l, r):
    """f(r) = beta(r) / r and df/dr"""
    g = np.exp(-r * r / (2 * SIGMA**2))
    f = r**l * g
    df = (l * r ** max(l - 1, 0) if l else 0.0) * g - r ** (l + 1) / SIGMA**2 * g
    return f, df


def _block(vals, per_line=4):
    lines = []
    for i in range(0, len(vals), per_line):
        lines.append("  ".join(f"{v: .15E}" for v in vals[i:i + per_line]))
    return "\n".join(lines)


def write_upf(path, projectors, energies, z_valence=6.0):
    """projectors: [(l, cutoff_radius)], energies in Ha (written as D_kk in Ry)"""
    r = np.arange(MESH_N) * MESH_DR
    n = len(projectors)
    parts = [f'<UPF version="2.0.1">',
             '<PP_INFO>synthetic Gaussian projectors</PP_INFO>',
             f'<PP_HEADER generated="test" element="X" pseudo_type="NC" z_valence="{z_valence:.2f}" '
             f'mesh_size="{MESH_N}" number_of_proj="{n}"/>',
             '<PP_MESH>',
             f'<PP_R type="real" size="{MESH_N}" columns="4">', _block(r), '</PP_R>',
             '</PP_MESH>',
             '<PP_NONLOCAL>']
    for k, (l, rc) in enumerate(projectors):
        beta = r ** (l + 1) * np.exp(-r * r / (2 * SIGMA**2))
        parts += [f'<PP_BETA.{k + 1} type="real" size="{MESH_N}" columns="4" index="{k + 1}" '
                  f'angular_momentum="{l}" cutoff_radius_index="{int(round(rc / MESH_DR))}" '
                  f'cutoff_radius="{rc:.10E}">', _block(beta), f'</PP_BETA.{k + 1}>']
    d = np.zeros((n, n))
    d[np.arange(n), np.arange(n)] = 2.0 * np.asarray(energies, dtype=np.float64)
    parts += [f'<PP_DIJ type="real" size="{n * n}" columns="4">', _block(d.ravel()), '</PP_DIJ>',
              '</PP_NONLOCAL>', '</UPF>', '']
    with open(path, "w") as fh:
        fh.write("\n".join(parts))
    return str(path)
