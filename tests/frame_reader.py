"""An independent reader of binary trajectory frames, restating the reference's
reader: Simul::readObjects (simul_file.cc:240-390), Object::readReference
(object.cc:109-152), FiberNaked::read (fiber_naked.cc:1208-1240),
PointSet::read (point_set.cc:346-366), Solid::read (solid.cc:908-925) and
InputWrapper's binary reads (iowrapper.cc:42-283).  Frames whose #time line
ends in ", real 8" hold 8-byte reals (the lossless form); the reference reads
4-byte floats.  Test infrastructure: it checks the frames the oracle and the
device write."""
import re
import struct

import numpy as np


class Frame:
    def __init__(self):
        self.time = None
        self.dim = None
        self.format = None
        self.real = 4
        self.binary = False
        self.objects = []      # dicts: tag, prop, number, mark, and the record's fields
        self.sections = []     # section names in order
        self.end = None        # offset after the frame


class _In:
    def __init__(self, data, pos=0):
        self.d = data
        self.p = pos

    def getc(self):
        if self.p >= len(self.d):
            return None
        c = self.d[self.p]
        self.p += 1
        return chr(c)

    def take(self, n):
        if self.p + n > len(self.d):
            raise ValueError("frame truncated")
        b = self.d[self.p:self.p + n]
        self.p += n
        return b

    def line(self):
        """FileWrapper::get_line: up to the separator, which is consumed"""
        k = self.d.find(b"\n", self.p)
        if k < 0:
            k = len(self.d)
        s = self.d[self.p:k].decode("latin-1")
        self.p = min(k + 1, len(self.d))
        return s


def _real(inp, fr):
    return struct.unpack("<d", inp.take(8))[0] if fr.real == 8 else float(struct.unpack("<f", inp.take(4))[0])


def _vector(inp, fr, dim):
    """InputWrapper::readFloatVector(double*, D): the file's coordinates, the
    missing ones 0, the extra ones read and dropped"""
    vals = [_real(inp, fr) for _ in range(fr.dim)]
    return (vals + [0.0] * dim)[:dim]


def _reference(inp, pretag):
    if pretag == "$":
        prop = struct.unpack("<H", inp.take(2))[0]
        number = struct.unpack("<I", inp.take(4))[0]
        mark = struct.unpack("<i", inp.take(4))[0]
    else:
        prop = struct.unpack("<B", inp.take(1))[0]
        number = struct.unpack("<H", inp.take(2))[0]
        mark = 0
    if number == 0:
        raise ValueError("Invalid (null) object reference")
    return prop, number, mark


def read_frame(data, dim, pos=0):
    """the first frame of `data` from `pos`; objects of the fiber and solid
    sections are parsed, other sections skipped to the next '#' line"""
    fr = Frame()
    inp = _In(data, pos)
    c = "\n"
    section = None
    started = False
    while True:
        tag = inp.getc()
        while tag is not None and tag.isspace():
            c = tag
            tag = inp.getc()
        if tag is None:
            break
        if c != "\n":
            inp.line()
            c = "\n"
            continue
        pretag = None
        if tag == "$":
            pretag, tag = tag, inp.getc()
        if tag == "#":
            line = inp.line()
            if line.startswith("Cytosim "):
                started = True
            elif line.startswith("section "):
                section = line[8:]
                fr.sections.append(section)
            elif line.startswith("time "):
                m = re.match(r"time ([^,]+), dim (\d+), format (\d+)(.*)", line)
                fr.time, fr.dim, fr.format = float(m.group(1)), int(m.group(2)), int(m.group(3))
                fr.real = 8 if m.group(4) == ", real 8" else 4
            elif line.startswith("binary "):
                if line[7:9] != "01":
                    raise ValueError("byte order not handled by this reader")
                fr.binary = True
            elif line.startswith("end cytosim "):
                fr.end = inp.p
                return fr
            continue
        if not fr.binary:
            raise ValueError("only binary frames are read here")
        if section == "fiber" and tag == "f":
            prop, number, mark = _reference(inp, pretag)
            sig = struct.unpack("<I", inp.take(4))[0]
            length, cut, absc = _real(inp, fr), _real(inp, fr), _real(inp, fr)
            n = struct.unpack("<H", inp.take(2))[0]
            pts = [_vector(inp, fr, dim) for _ in range(n)]
            fr.objects.append(dict(tag="f", prop=prop, number=number, mark=mark, signature=sig, length=length,
                                   cut_wished=cut, abscissa=absc, points=np.array(pts), radii=np.zeros(n)))
        elif section == "solid" and tag == "d":
            prop, number, mark = _reference(inp, pretag)
            n = struct.unpack("<H", inp.take(2))[0]
            pts, rad = [], []
            for _ in range(n):
                pts.append(_vector(inp, fr, dim))
                rad.append(_real(inp, fr))
            fr.objects.append(dict(tag="d", prop=prop, number=number, mark=mark, points=np.array(pts),
                                   radii=np.array(rad)))
        else:
            # a record of another section: skip to the next line that opens with '#'
            k = data.find(b"\n#", inp.p)
            inp.p = len(data) if k < 0 else k
            c = "\n"
    if not started:
        raise ValueError("no frame")
    fr.end = inp.p
    return fr
