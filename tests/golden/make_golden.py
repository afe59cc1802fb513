"""Write the golden vectors of the kernels that no closed form pins.

The HLLD sweep's star and double-star branches with the limiters and the
first-order fallback, the CT-contact upwind weights, the face averages of the
resistive energy flux and the limited transverse gradient of the anisotropic heat
flux are not fixed by any closed form or invariant that the reference's tests hold.
This script runs the reference's own scalar kernels --
``kernels/mhd_sweep/mhd_sweep_pyccel.py``, ``kernels/ct_contact/ct_contact_pyccel.py``,
``kernels/mhd_resistive/mhd_resistive_pyccel.py`` and
``kernels/mhd_conductive/mhd_conductive_pyccel.py``, which import only ``math`` and
``numpy`` -- as plain Python on the seeded blocks of the reference's kernel test
(``tests/unit/test_mhd_kernels.py:17-33,54-76``) and stores inputs and outputs in
``golden.npz`` next to this file.  Run it only where the reference tree exists:

    python tests/golden/make_golden.py /root/reference
"""

from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np

SHAPE = (13, 9)
GAMMA = 5.0 / 3.0


def _load(path: Path, name: str):
    spec = importlib.util.spec_from_file_location(name, path)
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def block(seed=1):
    """The seeded block of test_mhd_kernels.py:17-33."""
    rng = np.random.default_rng(seed)
    nx, ny = SHAPE
    prim = rng.normal(size=(8, nx, ny))
    prim[0] = rng.uniform(0.5, 2.0, SHAPE)
    prim[4] = rng.uniform(0.1, 1.0, SHAPE)
    prim[1, :, 6] += 30.0
    prim[2, 8, :] -= 30.0
    prim[4, 5, 3] = 1.0e-3
    bxf = rng.normal(size=(nx + 1, ny))
    byf = rng.normal(size=(nx, ny + 1))
    bxf[4, 2] = byf[2, 4] = 0.0
    return prim, bxf, byf


def ct_inputs(seed=2):
    """The seeded fluxes of test_mhd_kernels.py:66-76."""
    rng = np.random.default_rng(seed)
    nx, ny = SHAPE
    fx = rng.normal(size=(7, nx - 3, ny))
    fy = rng.normal(size=(7, nx, ny - 3))
    fx[0, 3, 4] = fy[0, 2, 5] = 0.0
    ecc = rng.normal(size=SHAPE)
    return fx, fy, ecc


#: Coefficients of the non-ideal vectors: dx, dy, eta, parallel, perpendicular.
NONIDEAL = (0.1, 0.07, 0.37, 0.8, 0.05)


def nonideal_inputs(seed=3):
    """Face fluxes and node E_z to add the non-ideal terms to (any values do)."""
    rng = np.random.default_rng(seed)
    nx, ny = SHAPE
    fx = rng.normal(size=(7, nx - 3, ny))
    fy = rng.normal(size=(7, nx, ny - 3))
    ez = rng.normal(size=(nx - 3, ny - 3))
    return fx, fy, ez


def field_free_face(prim):
    """The block with no field at the x-face 4 of row 2 (bxf[4, 2] = 0 already) nor at
    the y-face 4 of column 2 (byf[2, 4] = 0), where the unit field is 0 / tiny."""
    prim = prim.copy()
    prim[5:8, 3:5, 2] = 0.0
    prim[5:8, 2, 3:5] = 0.0
    return prim


def rotor_column_block():
    """Four x-columns (32..35, haloed) of the stage-2 input of step 17 of the rotor of
    ``mhd.py:561-592`` at 32^2 (this repository's oracle run): at rest far from the
    rotor, so a y-face has B_n at round-off and its tangential fluxes cancel to 1e-15,
    where a single rounding of the HLLD shows."""
    repo = Path(__file__).resolve().parents[2]
    sys.path.insert(0, str(repo))
    from oracle import kernels as K
    from tests.helpers import oracle2d
    from wellbalanced_amd import initial

    o = oracle2d(initial.rotor(32))
    for _ in range(16):
        o.step(o.stable_step())
    o.advance(o.stable_step())
    o.fill_halos()
    prim, _, byf, _, _ = K.primitive(*o.cons, o.az, o.dx, o.dy, *o.b0, o.gamma)
    return np.ascontiguousarray(prim[:, 32:37, :]), np.ascontiguousarray(byf[32:37, :]), o.gamma


def main(reference: Path) -> None:
    kernels = reference / "src" / "carafe" / "kernels"
    sweep = _load(kernels / "mhd_sweep" / "mhd_sweep_pyccel.py", "ref_mhd_sweep")
    ct = _load(kernels / "ct_contact" / "ct_contact_pyccel.py", "ref_ct_contact")
    out: dict[str, np.ndarray] = {}
    nx, ny = SHAPE
    for seed in (1, 7):
        prim, bxf, byf = block(seed)
        out[f"prim_{seed}"] = prim
        out[f"bxf_{seed}"] = bxf
        out[f"byf_{seed}"] = byf
        for limiter in range(4):
            for axis in (0, 1):
                shape = (7, nx - 3, ny) if axis == 0 else (7, nx, ny - 3)
                flux = np.full(shape, np.nan)
                sweep.mhd_sweep(prim, (bxf, byf)[axis], axis, limiter, GAMMA, flux)
                out[f"flux_{seed}_{axis}_{limiter}"] = flux
    prim, byf, gamma = rotor_column_block()
    flux = np.full((7, prim.shape[1], prim.shape[2] - 3), np.nan)
    for limiter in range(4):
        sweep.mhd_sweep(prim, byf, 1, limiter, gamma, flux)
        out[f"rotor_flux_{limiter}"] = flux.copy()
    out.update(rotor_prim=prim, rotor_byf=byf, rotor_gamma=np.array(gamma))
    fx, fy, ecc = ct_inputs()
    ez = np.full((nx - 3, ny - 3), np.nan)
    ct.ct_contact(fx[5], fy[5], fx[0], fy[0], ecc, ez)
    out.update(ct_fx=fx, ct_fy=fy, ct_ecc=ecc, ct_ez=ez)
    res = _load(kernels / "mhd_resistive" / "mhd_resistive_pyccel.py", "ref_mhd_resistive")
    con = _load(kernels / "mhd_conductive" / "mhd_conductive_pyccel.py", "ref_mhd_conductive")
    dx, dy, eta, parallel, perpendicular = NONIDEAL
    fx0, fy0, ez0 = nonideal_inputs()
    out.update(ni_fx=fx0, ni_fy=fy0, ni_ez=ez0)
    for seed in (1, 7):
        prim, bxf, byf = block(seed)
        fx, fy, ez = fx0.copy(), fy0.copy(), ez0.copy()
        res.mhd_resistive(prim, bxf, byf, dx, dy, eta, fx, fy, ez)
        out.update({f"res_fx_{seed}": fx, f"res_fy_{seed}": fy, f"res_ez_{seed}": ez})
        cprim = field_free_face(prim)
        fx, fy = fx0.copy(), fy0.copy()
        con.mhd_conductive(cprim, bxf, byf, dx, dy, parallel, perpendicular, fx, fy)
        out.update({f"con_prim_{seed}": cprim, f"con_fx_{seed}": fx, f"con_fy_{seed}": fy})
    target = Path(__file__).resolve().parent / "golden.npz"
    np.savez_compressed(target, **out)
    print(f"wrote {target} ({len(out)} arrays)")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
