"""Pins the oracle's trajectory frame (Simul::writeObjects in binary mode,
simul_file.cc:402-489) before the device frame is compared with it.

* Read back by an independent restatement of the reference's reader
  (tests/frame_reader.py), a frame gives every record's identity, property,
  mark and fields, and the points rounded to float32 (numpy's conversion); the
  fiber length equals the float32 of the polyline length of straight fibers.
* The lossless form gives back the FP64 points and time bit for bit.
* Object::writeReference switches to the long form ('$') exactly past number
  65535, property 255, or at a nonzero mark.
* A two-point fiber in 2D gives the literal bytes of the reference's writer.
* Sections handed in as bytes land before the fibers and after the solids, and
  two frames written back to back read as two frames.
"""
import struct

import numpy as np
import pytest

from frame_reader import read_frame

PARAMS = dict(time_step=0.005, kT=0.0042, viscosity=0.02)
DATE = "Fri Dec  4 10:11:12 2026"


def _scene(oracle, dim, seed, nfib=7, nsol=3):
    """straight fibers (segments of exactly their segmentation up to rounding)
    and Solids; returns the oracle, records and the points per Mecable"""
    rng = np.random.default_rng(seed)
    m = oracle.OracleMeca(dim, PARAMS)
    pts, recs, rads = [], [], []
    for f in range(nfib):
        n = int(rng.integers(2, 12))
        cut = float(rng.uniform(0.1, 0.9))
        d = rng.standard_normal(dim)
        d /= np.linalg.norm(d)
        p = rng.uniform(-3, 3, dim) + cut * np.arange(n)[:, None] * d
        m.add_fiber(n, cut, 20.0)
        pts.append(p)
        rads.append(np.zeros(n))
        recs.append((int(rng.integers(1, 200)), 0, int(rng.integers(0, 4)), int(rng.integers(0, 2**32)),
                     float(rng.uniform(0.1, 1)), float(rng.uniform(-5, 5))))
    for k in range(nsol):
        n = int(rng.integers(2, 8))
        r = rng.uniform(0.01, 0.3, n)
        p = rng.uniform(-2, 2, dim) + 0.5 * rng.standard_normal((n, dim))
        m.add_solid(r, p.ravel())
        pts.append(p)
        rads.append(r)
        recs.append((int(rng.integers(1, 200)), 0, int(rng.integers(0, 3)), 0, 0.0, 0.0))
    m.prepare(np.concatenate([p.ravel() for p in pts]))
    return m, recs, pts, rads, nfib


@pytest.mark.parametrize("dim", [2, 3])
def test_float_frame_reads_back(oracle, dim):
    m, recs, pts, rads, nfib = _scene(oracle, dim, 3 + dim)
    data = m.write_frame(recs, 12.25, DATE, 4)
    fr = read_frame(data, dim)
    assert data[fr.end:] == b"\n"   # the frame's last newline, before the next frame
    assert fr.time == 12.25 and fr.dim == dim and fr.format == 41 and fr.real == 4
    assert fr.sections == ["fiber", "solid", "end"]
    assert len(fr.objects) == len(recs)
    for k, (o, r, p, rd) in enumerate(zip(fr.objects, recs, pts, rads)):
        assert o["tag"] == ("f" if k < nfib else "d")
        assert (o["number"], o["mark"], o["prop"]) == (r[0], r[1], r[2])
        assert np.array_equal(o["points"], p.astype(np.float32).astype(np.float64))
        assert np.array_equal(o["radii"], rd.astype(np.float32).astype(np.float64))
        if k < nfib:
            assert o["signature"] == r[3]
            assert o["cut_wished"] == np.float32(r[4]) and o["abscissa"] == np.float32(r[5])
            polyline = np.linalg.norm(np.diff(p, axis=0), axis=1).sum()
            assert abs(o["length"] - polyline) <= 2 * np.spacing(np.float32(polyline))


@pytest.mark.parametrize("dim", [2, 3])
def test_lossless_frame_is_exact(oracle, dim):
    m, recs, pts, rads, nfib = _scene(oracle, dim, 11 + dim)
    t = 0.1 + 0.2   # not a short decimal
    data = m.write_frame(recs, t, DATE, 8)
    fr = read_frame(data, dim)
    assert fr.real == 8 and fr.time == t
    for k, (o, r, p, rd) in enumerate(zip(fr.objects, recs, pts, rads)):
        assert np.array_equal(o["points"], p)
        assert np.array_equal(o["radii"], rd)
        if k < nfib:
            assert o["cut_wished"] == r[4] and o["abscissa"] == r[5]


@pytest.mark.parametrize("number,prop,mark,long_form", [
    (65535, 255, 0, False), (65536, 255, 0, True), (65535, 256, 0, True), (7, 1, -3, True), (1, 0, 0, False)])
def test_reference_form(oracle, number, prop, mark, long_form):
    m = oracle.OracleMeca(3, PARAMS)
    m.add_fiber(2, 0.5, 20.0)
    m.prepare(np.array([0, 0, 0, 0.5, 0, 0], float))
    data = m.write_frame([(number, mark, prop, 9, 0.5, 0.0)], 0.0, DATE, 4)
    k = data.index(b"#section fiber\n") + len(b"#section fiber\n")
    assert (data[k:k + 1] == b"$") == long_form
    o = read_frame(data, 3).objects[0]
    assert (o["number"], o["prop"], o["mark"]) == (number, prop, mark)


def test_known_bytes_of_a_small_frame(oracle):
    m = oracle.OracleMeca(2, PARAMS)
    m.add_fiber(2, 0.75, 20.0)
    m.prepare(np.array([1.0, -2.0, 1.75, -2.0]))
    data = m.write_frame([(7, 0, 1, 0x01020304, 0.5, -1.25)], 1.5, DATE, 4)
    want = (b"\n\n#Cytosim  Fri Dec  4 10:11:12 2026"
            b"\n#time 1.500000, dim 2, format 41"
            b"\n#binary 01\n"
            b"\n#section fiber"
            b"\nf\x01\x07\x00" b"\x04\x03\x02\x01"
            + struct.pack("<fff", 0.75, 0.5, -1.25) + b"\x02\x00"
            + struct.pack("<ffff", 1.0, -2.0, 1.75, -2.0)
            + b"\n#section end"
            b"\n#end cytosim Fri Dec  4 10:11:12 2026\n\n")
    assert data == want


def test_sections_around_and_frames_back_to_back(oracle):
    m, recs, pts, rads, nfib = _scene(oracle, 3, 21, nfib=3, nsol=2)
    before = b"\n#section space\ne\x01\x01\x00"   # bytes the host's own writers made
    after = b"\n#section couple\nc\x01\x02\x00"
    a = m.write_frame(recs, 1.0, DATE, 4, before, after)
    assert a.index(before) < a.index(b"\n#section fiber") < a.index(b"\n#section solid") < a.index(after)
    assert a.index(after) < a.index(b"\n#section end")
    b = m.write_frame(recs, 2.0, DATE, 8)
    both = a + b
    f1 = read_frame(both, 3)
    f2 = read_frame(both, 3, f1.end)
    assert f1.sections == ["space", "fiber", "solid", "couple", "end"]
    assert (f1.time, f2.time, f2.end) == (1.0, 2.0, len(both) - 1)
    assert len(f1.objects) == len(f2.objects) == len(recs)
    assert np.array_equal(f2.objects[0]["points"], pts[0])


def test_dimension_conversion_of_the_reader():
    """readFloatVector: a 2D frame read in 3D gets z = 0; a 3D frame read in 2D
    drops z (iowrapper.cc:261-287) -- pins the test reader itself"""
    body = (b"\n\n#Cytosim  x\n#time 0.000000, dim 3, format 41\n#binary 01\n\n#section fiber\nf\x00\x01\x00"
            + b"\x00" * 4 + struct.pack("<fff", 1, 1, 0) + b"\x02\x00"
            + struct.pack("<ffffff", 1, 2, 3, 4, 5, 6) + b"\n#section end\n#end cytosim x\n\n")
    o = read_frame(body, 2).objects[0]
    assert np.array_equal(o["points"], [[1, 2], [4, 5]])
    o3 = read_frame(body.replace(b"dim 3", b"dim 2")[:-len(b"\n#section end\n#end cytosim x\n\n")]
                    .replace(struct.pack("<ffffff", 1, 2, 3, 4, 5, 6), struct.pack("<ffff", 1, 2, 4, 5))
                    + b"\n#section end\n#end cytosim x\n\n", 3).objects[0]
    assert np.array_equal(o3["points"], [[1, 2, 0], [4, 5, 0]])
