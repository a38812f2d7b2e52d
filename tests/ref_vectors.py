"""Test vectors in the layout the reference's test driver reads (main_gtest.cpp:64-76,
140-149): input.npy, input_tiled_p4_ref.npy, fop_ref.npy, thresholds.npy,
det_loc_ref.npy, det_pwr_ref.npy in one directory, produced the way ft_conv.py and
hsum.py produce them (ft_conv.py:123-207, hsum.py:91-151), here from the oracle:
FOP in FP64 rounded to float32, harmonic planes summed in float32, the mock
quantile thresholds (about 64 candidates per plane) and every detection above
them (uncapped), encoded with gen_info.h:62-63."""

from __future__ import annotations

import shutil
from pathlib import Path

import numpy as np

from oracle import fdas_oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


def write_vectors(vec_dir: Path, x: np.ndarray, bank: np.ndarray, n_harmonics: int = 8, n_cands: int = 64,
                  tile: int = 2048) -> dict:
    vec_dir.mkdir(parents=True, exist_ok=True)
    n = x.size
    t_count = bank.shape[0]
    c = t_count // 2
    geo = O.geometry(n, tile, bank.shape[1])
    tiles = O.to_fft_order(O.tiles_transformed(x, geo)).astype(np.complex64)
    fop = O.fop_fft(x, bank).astype(np.float32)
    hps = O.harmonic_planes(fop, n_harmonics, dtype=np.float32).astype(np.float32)
    thr = O.quantile_thresholds(hps, n_cands)
    loc, pwr = [], []
    for h in range(n_harmonics):
        rr, ff = np.nonzero(hps[h] > thr[h])
        loc.append(O.encode_location(h + 1, rr - c, ff, c))
        pwr.append(hps[h][rr, ff])
    loc = np.concatenate(loc).astype(np.uint32)
    pwr = np.concatenate(pwr).astype(np.float32)
    np.save(vec_dir / "input.npy", np.ascontiguousarray(x, dtype=np.complex64))
    np.save(vec_dir / "input_tiled_p4_ref.npy", tiles)
    np.save(vec_dir / "fop_ref.npy", fop)
    np.save(vec_dir / "thresholds.npy", thr)
    np.save(vec_dir / "det_loc_ref.npy", loc)
    np.save(vec_dir / "det_pwr_ref.npy", pwr)
    return {"n_detections": int(loc.size), "thresholds": thr}


def driver_tree(root: Path) -> Path:
    """The directory layout the reference test driver expects: it reads the template file from
    ../../fdas/tmpl/ relative to its working directory (main_gtest.cpp:503).
    Returns that working directory."""
    tmpl = root / "fdas" / "tmpl"
    tmpl.mkdir(parents=True, exist_ok=True)
    shutil.copyfile(GOLDEN / "fdas_templates_85_350_ft_p4.npy", tmpl / "fdas_templates_85_350_ft_p4.npy")
    cwd = root / "build" / "run"
    cwd.mkdir(parents=True, exist_ok=True)
    return cwd
