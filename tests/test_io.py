"""CPU tests of the front-end reader and the candidate wire format."""

import struct

import numpy as np
import pytest

from fdas_amd import candidates as K
from fdas_amd.io import FrontEndError, read_tim, trim_series
from oracle import fdas_oracle as O


def write_tim(path, samples, tsamp=6.4e-5, data_type=2, nbits=32, end=True):
    """A SIGPROC time series: length-prefixed keywords with their values, then float32 samples."""
    def key(k):
        return struct.pack("<i", len(k)) + k.encode()
    hdr = key("HEADER_START") + key("source_name") + key("J0000+00") + key("data_type") + struct.pack("<i", data_type)
    hdr += key("nchans") + struct.pack("<i", 1) + key("nbits") + struct.pack("<i", nbits)
    hdr += key("tsamp") + struct.pack("<d", tsamp) + key("refdm") + struct.pack("<d", 12.5)
    if end:
        hdr += key("HEADER_END")
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.asarray(samples, np.float32).tobytes())


def test_read_tim_with_header(tmp_path):
    x = np.random.default_rng(1).standard_normal(1000).astype(np.float32)
    p = tmp_path / "a.tim"
    write_tim(p, x, tsamp=1.28e-4)
    s, ts = read_tim(p)
    assert np.array_equal(s, x) and ts == 1.28e-4


def test_read_tim_raw_and_header_skipped(tmp_path):
    x = np.arange(64, dtype=np.float32)
    p = tmp_path / "raw.tim"
    x.tofile(p)
    s, ts = read_tim(p)
    assert np.array_equal(s, x) and ts is None
    q = tmp_path / "h.tim"
    write_tim(q, x)
    s2, ts2 = read_tim(q, read_header=False)  # the header bytes are read as samples, as the reference does
    assert s2.size == q.stat().st_size // 4 and ts2 is None


@pytest.mark.parametrize("kw", [{"nbits": 8}, {"data_type": 1}, {"end": False}])
def test_read_tim_rejects_other_formats(tmp_path, kw):
    p = tmp_path / "bad.tim"
    write_tim(p, np.zeros(16, np.float32), **kw)
    with pytest.raises(FrontEndError):
        read_tim(p)


def test_read_tim_missing_file(tmp_path):
    with pytest.raises(FrontEndError, match="cannot open"):
        read_tim(tmp_path / "none.tim")


def test_trim_series():
    x = np.arange(10, dtype=np.float32)
    assert np.array_equal(trim_series(x, 4), x[:8])
    with pytest.raises(ValueError):
        trim_series(x, 6)


def test_location_encoding_known_answers():
    assert int(K.invalid_location(85)) == 0x15400000  # gen_info.h:69, encode(1, c + 1, 0)
    assert int(K.encode(8, 42, (1 << 22) - 1, 85)) == 0xF53FFFFF
    h, t, f = K.decode(np.array([0x15400000, 0xF53FFFFF], np.uint32), 85)
    assert list(h) == [1, 8] and list(t) == [43, 42] and list(f) == [0, (1 << 22) - 1]
    assert list(K.valid(np.array([0x15400000, 0xF53FFFFF], np.uint32), 85)) == [False, True]


def test_encoding_matches_oracle():
    rng = np.random.default_rng(3)
    k = rng.integers(1, 9, 500)
    t = rng.integers(-42, 43, 500)
    f = rng.integers(0, 1 << 22, 500)
    assert np.array_equal(K.encode(k, t, f, 85), O.encode_location(k, t, f, 42))
    h, tt, ff = K.decode(K.encode(k, t, f, 85), 85)
    assert np.array_equal(h, k) and np.array_equal(tt, t) and np.array_equal(ff, f)


class _Result:
    def __init__(self, keys, powers):
        self.keys, self.powers = keys, powers
        self.n_detected = np.array([len(k) for k in keys])


def test_lists_from_result_and_table(tmp_path):
    c = 42
    keys = [np.array([(c << 22) | 7, ((c + 3) << 22) | 100], np.uint32), np.array([((c - 5) << 22) | 9], np.uint32)]
    pw = [np.array([5.0, 6.0], np.float32), np.array([9.5], np.float32)]
    loc, pwr = K.lists_from_result(_Result(keys, pw), 85, 4)
    assert loc.size == 8 and pwr.size == 8
    assert list(K.decode(loc[:2], 85)[1]) == [0, 3] and list(pwr[2:4]) == [-1.0, -1.0]
    assert int(loc[4]) == int(K.encode(2, -5, 9, 85)) and int(loc[5]) == 0x15400000
    lines = K.format_table(loc, pwr, 85)
    assert lines[3:] == ["[   1] 1   0       7 5.000000", "[   2] 1   3     100 6.000000",
                         "[   5] 2  -5       9 9.500000"]
    K.save_lists(tmp_path / "loc.npy", tmp_path / "pwr.npy", loc, pwr)
    l2, p2 = K.load_lists(tmp_path / "loc.npy", tmp_path / "pwr.npy")
    assert np.array_equal(l2, loc) and np.array_equal(p2, pwr)
    freqs = K.bin_frequencies(1 << 10, 1e-3)
    assert np.allclose(freqs, np.fft.fftfreq(2048, 1e-3)[:1024].astype(np.float32))
    np.save(tmp_path / "f.npy", freqs)
    assert K.main([str(tmp_path / "loc.npy"), str(tmp_path / "pwr.npy"), "-F", str(tmp_path / "f.npy")]) == 0
    with pytest.raises(ValueError):
        K.lists_from_result(_Result(keys, pw), 169, 4)
