"""Candidate wire format: the reference's 32-bit location lists, their files and decoder.

A location packs ``((k - 1) & 7) << 29 | ((t + c) & 0x7f) << 22 | (f & 0x3fffff)``
(``gen_info.h:62-67``); a list holds ``cands_per_harmonic`` entries per plane,
plane-major, unused entries set to ``encode(1, c + 1, 0)`` with power -1.0
(``gen_info.h:69``, ``detect.cl.mako:38-39``).  The files are the reference's
``.npy`` pair (uint32 locations, float32 powers) that ``print_candidates.py``
reads; ``python -m fdas_amd.candidates LOC PWR [-F FREQS] [-N T]`` prints them
in the same table form (``print_candidates.py:56-80``) without the
``bitstring`` dependency.  ``lists_from_result`` turns an engine result into
that layout; ``bin_frequencies`` is the reference's ``fftfreq`` column
(``ft_conv.py:134,145``).
"""

from __future__ import annotations

import argparse
import sys

import numpy as np

BIN_BITS = 22
INVALID_POWER = np.float32(-1.0)


def encode(harmonic, template, freq_bin, n_templates=85):
    """32-bit location of (harmonic k >= 1, template t in [-c, c], bin f)."""
    c = n_templates // 2
    k = np.asarray(harmonic, np.int64)
    t = np.asarray(template, np.int64)
    f = np.asarray(freq_bin, np.int64)
    return ((((k - 1) & 7) << 29) | (((t + c) & 0x7F) << 22) | (f & 0x3FFFFF)).astype(np.uint32)


def invalid_location(n_templates=85):
    c = n_templates // 2
    return encode(1, c + 1, 0, n_templates)


def decode(locations, n_templates=85):
    """(harmonic, template, bin) arrays of 32-bit locations."""
    loc = np.asarray(locations, np.uint32).astype(np.int64)
    c = n_templates // 2
    return ((loc >> 29) & 7) + 1, ((loc >> 22) & 0x7F) - c, loc & 0x3FFFFF


def valid(locations, n_templates=85):
    """False for the empty-slot marker (harmonic 1 with a template past c; print_candidates.py:75-76)."""
    h, t, _ = decode(locations, n_templates)
    return ~((h == 1) & (t > n_templates // 2))


def lists_from_result(result, n_templates, cands_per_harmonic, n_harmonics=None):
    """Reference-layout (loc, pwr) arrays of an engine ``TrialResult``: plane-major,
    ``cands_per_harmonic`` entries per plane, sorted by (template, bin), padded.
    Only T <= 127 and H <= 8 fit the 7- and 3-bit fields."""
    H = result.n_detected.size if n_harmonics is None else n_harmonics
    if n_templates > 127 or H > 8:
        raise ValueError("T > 127 or H > 8 do not fit the 32-bit candidate layout")
    cap = cands_per_harmonic
    c = n_templates // 2
    loc = np.full(H * cap, invalid_location(n_templates), np.uint32)
    pwr = np.full(H * cap, INVALID_POWER, np.float32)
    for k in range(H):
        keys = np.asarray(result.keys[k], np.uint32)[:cap]
        rows = (keys >> BIN_BITS).astype(np.int64)
        bins = (keys & ((1 << BIN_BITS) - 1)).astype(np.int64)
        m = keys.size
        loc[k * cap:k * cap + m] = encode(k + 1, rows - c, bins, n_templates)
        pwr[k * cap:k * cap + m] = np.asarray(result.powers[k], np.float32)[:cap]
    return loc, pwr


def save_lists(loc_path, pwr_path, loc, pwr):
    """The reference's two ``.npy`` files (uint32 locations, float32 powers)."""
    loc = np.asarray(loc)
    pwr = np.asarray(pwr)
    if loc.dtype != np.uint32 or pwr.dtype != np.float32 or loc.size != pwr.size:
        raise ValueError("locations must be uint32 and powers float32 of the same size")
    np.save(loc_path, loc)
    np.save(pwr_path, pwr)


def load_lists(loc_path, pwr_path):
    """Loads and checks a location / power pair (no pickles are read)."""
    loc = np.load(loc_path, allow_pickle=False)
    pwr = np.load(pwr_path, allow_pickle=False)
    if loc.dtype != np.uint32:
        raise ValueError(f"invalid data type in locations file '{loc_path}'")
    if pwr.dtype != np.float32:
        raise ValueError(f"invalid data type in power file '{pwr_path}'")
    if loc.size != pwr.size:
        raise ValueError(f"list size mismatch: {loc.size} vs. {pwr.size}")
    return loc, pwr


def bin_frequencies(n_bins, t_samp):
    """Frequency of bins 0..N-1 of an Ns = 2N point transform with sampling time t_samp, float32."""
    ns = 2 * n_bins
    return (np.arange(n_bins, dtype=np.float64) / (ns * t_samp)).astype(np.float32)


def format_table(loc, pwr, n_templates=85, frequencies=None):
    """Lines of the reference printer's table: one per valid entry, 1-based list index."""
    h, t, f = decode(loc, n_templates)
    ok = valid(loc, n_templates)
    if frequencies is not None:
        out = ["Harmonic  Template", "ID     │  │  Frequency    Power", "────── ┴ ─┴─ ──────────── ────────"]
    else:
        out = ["Harmonic  Template", "ID     │  │  Bin     Power", "────── ┴ ─┴─ ─────── ────────"]
    for i in np.nonzero(ok)[0]:
        if frequencies is not None:
            out.append(f"[{i + 1:4d}] {h[i]:1d} {t[i]:3d} {frequencies[f[i]]:12.6f} {pwr[i]:.6f}")
        else:
            out.append(f"[{i + 1:4d}] {h[i]:1d} {t[i]:3d} {f[i]:7d} {pwr[i]:.6f}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Decode 32-bit candidate lists and print them as a table.")
    ap.add_argument("locations", metavar="location-file", help=".npy file of uint32 encoded locations (harmonic, template, bin)")
    ap.add_argument("power", metavar="power-file", help=".npy file of float32 powers, one per location")
    ap.add_argument("-F", "--resolve-frequencies", dest="frequencies", metavar="frequencies-file",
                    help=".npy table mapping bin numbers to frequencies, printed beside each bin")
    ap.add_argument("-N", "--num-templates", dest="n_tmpl", metavar="n", type=int, default=85,
                    help="templates in the bank the lists came from (default 85)")
    args = ap.parse_args(argv)
    try:
        loc, pwr = load_lists(args.locations, args.power)
    except ValueError as e:
        print(f"[ERROR] {e}")
        return 1
    freqs = np.load(args.frequencies, allow_pickle=False) if args.frequencies else None
    for line in format_table(loc, pwr, args.n_tmpl, freqs):
        print(line)
    return 0


if __name__ == "__main__":
    sys.exit(main())
