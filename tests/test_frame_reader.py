"""cymg_read_frame, the library's reader of trajectory frames (host code, no
GPU), against the test reader that restates the reference's reader
(tests/frame_reader.py), on the oracle's frames: both forms, 2D and 3D, a 3D
frame read in 2D and a 2D frame read in 3D, sections of other objects around
the fibers and solids, frames back to back, and the errors of a truncated or a
text frame."""
import numpy as np
import pytest

from frame_reader import read_frame
from test_oracle_frame import DATE, _scene


def _same(got, fr, dim):
    assert got["time"] == fr.time
    assert len(got["objects"]) == len(fr.objects)
    for o, w in zip(got["objects"], fr.objects):
        assert (o["tag"], o["number"], o["prop"], o["mark"]) == (w["tag"], w["number"], w["prop"], w["mark"])
        k, n = o["first"], o["points"]
        assert n == len(w["points"])
        assert np.array_equal(got["points"][k:k + n].reshape(n, dim), w["points"])
        assert np.array_equal(got["radii"][k:k + n], w["radii"])
        if w["tag"] == "f":
            assert (o["signature"], o["length"], o["cut_wished"], o["abscissa"]) == \
                (w["signature"], w["length"], w["cut_wished"], w["abscissa"])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("rb", [4, 8])
@pytest.mark.parametrize("read_dim", [2, 3])
def test_reader_matches_the_reference_reader(pkg, oracle, dim, rb, read_dim):
    m, recs, pts, rads, nfib = _scene(oracle, dim, 40 + dim + rb)
    recs[1] = (70000, -2, 300, 5, 0.5, 1.0)      # long reference form
    data = m.write_frame(recs, 3.25, DATE, rb, b"\n#section space\ne\x01\x01\x00", b"\n#section couple\nc\x00\x02\x00")
    got = pkg.read_frame(data, read_dim)
    fr = read_frame(data, read_dim)
    _same(got, fr, read_dim)
    assert got["used"] == fr.end == len(data) - 1
    if rb == 8 and read_dim == dim:
        assert np.array_equal(got["points"], np.concatenate(pts))


def test_frames_back_to_back(pkg, oracle):
    m, recs, pts, rads, nfib = _scene(oracle, 3, 9)
    a = m.write_frame(recs, 1.0, DATE, 4)
    b = m.write_frame(recs, 2.0, DATE, 8)
    both = a + b
    f1 = pkg.read_frame(both, 3)
    f2 = pkg.read_frame(both[f1["used"]:], 3)
    assert (f1["time"], f2["time"]) == (1.0, 2.0)
    assert np.array_equal(f2["points"], np.concatenate(pts))


def test_reader_errors(pkg, oracle):
    m, recs, pts, rads, nfib = _scene(oracle, 3, 12)
    data = m.write_frame(recs, 1.0, DATE, 4)
    with pytest.raises(pkg.CymgError, match="truncated|end line"):
        pkg.read_frame(data[:len(data) // 2], 3)
    text = data[:data.index(b"\n#binary")] + b"\n#section fiber\nf1:7 0 1.0 0.5 0.0 2\n 0 0 0\n 1 0 0\n#section end\n"
    with pytest.raises(pkg.CymgError, match="binary"):
        pkg.read_frame(text, 3)
    with pytest.raises(pkg.CymgError, match="no frame"):
        pkg.read_frame(b"\n\n", 3)
