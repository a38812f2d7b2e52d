"""Writes the small golden fixtures the parity tests check the GPU path against.

Run from the repository root:  python tests/golden/make_golden.py
Inputs come from fdas_amd.synth (seeded); expected outputs from the FP64
oracle in oracle/fdas_oracle.py.  The two reference template banks in this
folder are data files copied from the reference repository
(fdas/tmpl/fdas_templates_85_350.npy, fdas/tmpl/fdas_templates_85_350_ft_p4.npy).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from fdas_amd import synth  # noqa: E402
from oracle import fdas_oracle as O  # noqa: E402


def small_case():
    """N = 3000 (ragged: not a multiple of the 1628-bin payload), 85 templates z = 2(j-42)."""
    n = 3000
    w = synth.Workload("golden", n, 11, [synth.Tone(900.0, 14.0, 4, 0.08)])
    x = w.spectrum()
    bank = O.make_bank(synth.drift_grid(85), 421)
    fop = O.fop_direct(x, bank)
    thr = O.noise_thresholds(x, 8)
    cands = O.select(O.detect(O.harmonic_planes(fop, 8), thr), 64)
    out = dict(spectrum=x, bank=bank, fop=fop.astype(np.float32), thresholds=thr,
               n_detected=np.array([c.n_detected for c in cands], dtype=np.int64))
    for h, c in enumerate(cands):
        out[f"rows_{h + 1}"] = c.rows
        out[f"bins_{h + 1}"] = c.bins
        out[f"powers_{h + 1}"] = c.powers
    np.savez_compressed(HERE / "golden_small.npz", **out)


def hsum_literal_case():
    """HSUM fixture from the literal per-element loop (hsum.py:93-120), float32, T = 9, N = 50."""
    rng = np.random.default_rng(7)
    fop = rng.exponential(1.0, size=(9, 50)).astype(np.float32)
    hps = O.harmonic_planes_loop(fop, 8)
    np.savez_compressed(HERE / "golden_hsum_tiny.npz", fop=fop, hps=hps)


if __name__ == "__main__":
    small_case()
    hsum_literal_case()
    for p in sorted(HERE.glob("*.npz")):
        print(p.name, p.stat().st_size)
