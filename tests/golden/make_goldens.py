"""Write golden fixtures from the compiled reference (oracle/_ref/libdftref.so).

Run in a container that has the reference checkout (``make -C oracle ref``
first).  The fixtures are inputs and the reference's outputs only; the tests
use them to pin the oracle and, on the GPU, the HIP path.
Usage: python tests/golden/make_goldens.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import Grid, Reference  # noqa: E402


def gaussians(g, centres, widths, weights):
    xs, ys, zs = g.coords()
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    out = np.zeros(g.shape)
    for (cx, cy, cz), a, w in zip(centres, widths, weights):
        out += w * np.exp(-a * ((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2))
    return out.ravel()


def main():
    R = Reference()
    rng = np.random.default_rng(20261203)
    out = {}

    # H psi on a non-cubic, anisotropic grid
    g = Grid(9, 8, 7, -3.1, 2.9, -2.0, 2.5, -1.7, 2.2)
    v = rng.standard_normal(g.size)
    psi = rng.standard_normal(g.size)
    out.update(hpsi_grid=np.array(g.n + g.bounds, dtype=float), hpsi_v=v, hpsi_psi=psi, hpsi_out=R.hpsi(g, v, psi))

    # Poisson (reference stop rule) on a 14^3 centred grid
    g = Grid.centered(14, 0.45)
    rho = gaussians(g, [(0.3, -0.2, 0.1), (-0.8, 0.5, -0.3)], [1.3, 0.9], [0.8, 0.5])
    vh = R.poisson(g, rho)
    out.update(poisson_grid=np.array(g.n + g.bounds, dtype=float), poisson_rho=rho, poisson_vh=vh,
               poisson_eh=np.array(R.hartree_energy(g, rho, vh)))

    # Poisson on a grid whose max-side faces miss xmax/ymax/zmax in FP64, so the
    # reference's boundary rule takes its Q/|r| branch (HartreePotentialMolecule.cpp:77-94)
    g = Grid(10, 9, 8, -3.3, 2.9, -2.7, 3.1, -1.9, 3.7)
    rho = gaussians(g, [(0.1, 0.2, 0.9)], [1.5], [1.0])
    out.update(poissonq_grid=np.array(g.n + g.bounds, dtype=float), poissonq_rho=rho,
               poissonq_vh=R.poisson(g, rho))

    # pointwise functionals on a lattice (zeta != 0, floor band, zero gradient, large gradient)
    na = np.array([1e-21, 4e-21, 1e-12, 1e-6, 1e-3, 0.05, 0.3, 2.0, 40.0])
    fr = np.array([0.0, 0.1, 0.5, 0.9, 1.0])
    gr = np.array([0.0, 1e-8, 0.3, 5.0])
    pts = []
    for n in na:
        for f in fr:
            for ga in gr:
                for gb in (0.0, 0.7):
                    pts.append([n * (1 - f), n * f, ga * n ** (4 / 3) + ga * 1e-3, gb * n])
    pts.append([-1e-3, 0.2, 0.1, 0.1])  # negative density is clamped
    pts = np.array(pts)
    out["xcpt_in"] = pts
    out["xcpt_pz81"] = R.xc_points(1, pts)
    out["xcpt_pbe96"] = R.xc_points(2, pts)

    # XC potential and energy on a grid with a floor region, spin-polarised
    g = Grid(11, 10, 9, -2.5, 2.5, -2.2, 2.3, -2.0, 2.0)
    na_ = gaussians(g, [(0.2, 0.1, -0.1)], [1.1], [0.9])
    nb_ = gaussians(g, [(-0.3, 0.2, 0.2)], [1.6], [0.4])
    na_[na_ < 2e-3] = 0.0
    nb_[nb_ < 1e-3] = 0.0
    out.update(xc_grid=np.array(g.n + g.bounds, dtype=float), xc_na=na_, xc_nb=nb_)
    for fid, name in ((1, "pz81"), (2, "pbe96")):
        out[f"xc_{name}_va"] = R.xc_potential(g, fid, na_, nb_, 0)
        out[f"xc_{name}_vb"] = R.xc_potential(g, fid, na_, nb_, 1)
        out[f"xc_{name}_exc"] = np.array(R.xc_energy(g, fid, na_, nb_))

    # density, kinetic (Neumann stencil), external, nuclear potential/repulsion
    g = Grid(8, 9, 10, -2.0, 2.0, -2.5, 2.0, -2.2, 2.6)
    psi = rng.standard_normal((3, g.size))
    el = np.array([1, 2, 0], dtype=np.int32)
    nuc = np.array([[2, 0.11, 0.23, 0.37], [1, 1.01, -0.49, 0.06], [8, -1.3, 0.4, -0.7]])
    vn = R.nuclear_potential(g, nuc)
    rho = np.abs(rng.standard_normal(g.size))
    out.update(misc_grid=np.array(g.n + g.bounds, dtype=float), misc_psi=psi, misc_el=el, misc_nuc=nuc,
               misc_density=R.spin_density(psi, el), misc_kinetic=np.array(R.kinetic_energy(g, psi, el)),
               misc_vnuc=vn, misc_rho=rho, misc_external=np.array(R.external_energy(g, rho, vn)),
               misc_enn=np.array(R.nuclear_repulsion(nuc)))

    # Davidson: 4 lowest states of a two-centre potential on 12^3
    g = Grid.centered(12, 0.5)
    nuc = np.array([[3, 0.25, 0.0, 0.1], [1, -1.25, 0.5, -0.3]])
    vn = R.nuclear_potential(g, nuc)
    eig, psi = R.davidson(g, vn, 4)
    out.update(dav_grid=np.array(g.n + g.bounds, dtype=float), dav_v=vn, dav_eig=eig, dav_psi=psi)

    # SCF known answers: main.cpp's H2 (7^3) and H2O (9^3), plus He on 20^3 h=0.5
    scf_cases = {
        "h2": (Grid(7, 7, 7, -8.1, 7.9, -8.1, 7.9, -8.1, 7.9), [[1, -0.7, 0, 0], [1, 0.7, 0, 0]], 1),
        "h2o": (Grid(9, 9, 9, -8.1, 7.9, -8.1, 7.9, -8.1, 7.9),
                [[8, 0, 0, 0], [1, 1.430, 0, 1.108], [1, -1.430, 0, 1.108]], 1),
        "he_pbe": (Grid.centered(16, 0.5), [[2, 0.0, 0.0, 0.0]], 2),
    }
    for name, (g, nuc, fid) in scf_cases.items():
        r = R.scf(g, nuc, fid)
        out[f"scf_{name}_grid"] = np.array(g.n + g.bounds, dtype=float)
        out[f"scf_{name}_nuc"] = np.array(nuc, dtype=float)
        out[f"scf_{name}_fid"] = np.array(fid)
        out[f"scf_{name}_result"] = np.array([r[k] for k in
                                              ("iterations", "T", "Eext", "EH", "Exc", "Enn", "total", "dE", "dRho", "ks")])
        out[f"scf_{name}_eig"] = r["eigenvalues"]
        print(name, r["total"], r["iterations"])

    # total density over both spin channels (mixed list order, an empty orbital) and its integral;
    # drawn last so the fixtures above keep their random inputs
    g = Grid(7, 8, 6, -1.9, 2.1, -2.0, 1.6, -1.4, 1.8)
    psi = rng.standard_normal((5, g.size))
    el = np.array([1, 1, 0, 2, 1], dtype=np.int32)
    spins = np.array([0, 1, 0, 1, 1], dtype=np.int32)
    dens = R.electron_density(psi, el, spins)
    out.update(edens_grid=np.array(g.n + g.bounds, dtype=float), edens_psi=psi, edens_el=el, edens_spins=spins,
               edens_density=dens, edens_integral=np.array(R.integrate_density(dens, g.dx, g.dy, g.dz)))

    path = os.path.join(HERE, "reference_goldens.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
