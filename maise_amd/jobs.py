"""The reference's file-driven relaxation job (JOBT 20/21: POSCAR in; OUTCAR, OSZICAR, CONTCAR out) on the GPU.

The reference's job (CELL_RELX, src/cmod.c:901-977, with READ_CELL, CELL_OUT and SAVE_CELL around it) reads the
POSCAR, evaluates and reports the start, minimises (GSL BFGS2 there; batched FIRE on the device here), reports
the final state and writes the CONTCAR.  Iteration 0 is the start, iteration 1 the relaxed state; the OUTCAR's
"optimizer type" line carries 4 for FIRE, a value the reference does not use for its own minimisers (0-3).
"""
from __future__ import annotations

import os

import numpy as np

from . import EV2GPA, Context, MaiseError, Model
from . import poscar as pc

MINT_FIRE = 4


def _species_index(model: Model, p: pc.Poscar, path: str) -> np.ndarray:
    """Model species index per atom from the POSCAR's symbols (the ATMN assignment of READ_CELL, src/util.c:933-951)."""
    idx = []
    for sym, n in zip(p.symbols, p.counts):
        z = pc.element_z(sym)
        if z not in model.spcz:
            raise MaiseError(f"species {z} in {path} is not among species in model file")
        idx += [model.spcz.index(z)] * n
    return np.array(idx, dtype=np.int32)


def relax_job(model: Model, poscar_path: str, out_dir: str, rlxt: int = 3, p_gpa: float = 0.0, nd: int = 3,
              max_steps: int = 3000, fmax: float = 0.01, smax_gpa: float = 0.01, detail: bool = True,
              version: str = "maise.3.0.00") -> dict:
    """Relax the structure of poscar_path: positions only (rlxt 2) or positions and cell at p_gpa (rlxt 3), to
    every |F| < fmax (eV/A) and, with the cell, every stress component < smax_gpa.  Appends iterations 0 and 1 to
    out_dir/OUTCAR and OSZICAR (detail: positions, forces, atom energies, stress) and writes out_dir/CONTCAR.
    Returns {"E", "H", "lattice", "x", "steps", "converged"}."""
    if rlxt not in (2, 3):
        raise ValueError("rlxt is 2 (positions) or 3 (positions and cell)")
    if rlxt == 3 and nd != 3:
        raise ValueError("a cell relaxation needs a periodic structure (nd = 3)")
    p = pc.read_poscar(poscar_path, nd=nd)
    if np.any(p.ff == 0):
        raise MaiseError(f"{poscar_path}: fixed coordinates (selective dynamics) are not supported by the device "
                         "relaxation")
    spc = _species_index(model, p, poscar_path)
    c = Context(model)
    try:
        c.set_structure(spc, p.lattice, p.x, nd=nd, xt=1)
        os.makedirs(out_dir, exist_ok=True)

        def report(it, lat):
            r = c.result(0)
            V = abs(np.linalg.det(lat))
            H = r["E"] + (p_gpa / EV2GPA) * V if nd == 3 else r["E"]
            pc.write_outcar(out_dir, it, lat, p.z, p.counts, r["E"], H, x=r["X"], F=r["F"], EA=r["EA"], U=r["U"],
                            version=version, model_id=model.id, mint=MINT_FIRE, rlxt=rlxt, p_gpa=p_gpa, nd=nd,
                            detail=detail)
            return r, H

        c.eval(forces=True)
        report(0, p.lattice)
        if rlxt == 2:
            steps, done = c.relax(max_steps=max_steps, fmax=fmax)
        else:
            steps, done = c.relax_cell(p_gpa=p_gpa, max_steps=max_steps, fmax=fmax, smax_gpa=smax_gpa)
        c.eval(forces=True)
        lat = c.lattice(0)
        r, H = report(1, lat)
        pc.write_poscar(os.path.join(out_dir, "CONTCAR"), lat, p.z, p.counts, r["X"], comment=p.comment, nd=nd,
                        ff=p.ff, velocities=p.velocities)
        return {"E": r["E"], "H": H, "lattice": lat, "x": r["X"], "steps": int(steps[0]), "converged": bool(done[0])}
    finally:
        c.close()
