"""Readers of the AMReX on-disk layouts the library writes (test helpers):
VisMF MultiFabs (<prefix>_H + data files), the checkpoint Header of
SaveCheckpoint (utilities.cpp:856-869) and the plotfile Header of
WriteSingleLevelPlotfile.  Written with numpy, independently of the library's
C++ reader, so a test that reads back what the library wrote checks the
format itself."""
from __future__ import annotations

import os
import re

import numpy as np

_BOX = re.compile(r"\(\((-?\d+),(-?\d+)\) \((-?\d+),(-?\d+)\) \((\d+),(\d+)\)\)")


def parse_boxes(text: str):
    """[(lo0, lo1, hi0, hi1, t0, t1)] of the first BoxArray "(n 0 ...)" in text"""
    m = re.search(r"\((\d+) 0\s*\n", text)
    assert m, "no BoxArray"
    nb = int(m.group(1))
    boxes = [tuple(int(x) for x in b.groups()) for b in _BOX.finditer(text, m.end())][:nb]
    assert len(boxes) == nb
    return boxes


def read_multifab(prefix: str):
    """VisMF MultiFab -> (header dict, [(box, array (ncomp, ny, nx))])"""
    text = open(prefix + "_H").read()
    lines = text.split("\n")
    hdr = {"version": int(lines[0]), "how": int(lines[1]), "ncomp": int(lines[2]), "ngrow": int(lines[3])}
    boxes = parse_boxes(text)
    fods = re.findall(r"FabOnDisk: (\S+) (\d+)", text)
    assert len(fods) == len(boxes)
    hdr["boxes"] = boxes
    mm = re.findall(r"^(\d+),(\d+)\n((?:[^\n]*,\n)+)", text.split("FabOnDisk")[-1], flags=re.M)
    hdr["minmax"] = [np.array([[float(x) for x in row.split(",")[:-1]] for row in blk.strip().split("\n")])
                     for _, _, blk in mm]
    out = []
    d = os.path.dirname(prefix)
    for box, (fname, off) in zip(boxes, fods):
        raw = open(os.path.join(d, fname), "rb").read()
        off = int(off)
        eol = raw.index(b"\n", off)
        head = raw[off:eol].decode()
        assert head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"), head
        fb = tuple(int(x) for x in _BOX.search(head).groups())
        assert fb == box
        nc = int(head.rsplit(" ", 1)[1])
        nx, ny = box[2] - box[0] + 1, box[3] - box[1] + 1
        a = np.frombuffer(raw, dtype="<f8", count=nc * nx * ny, offset=eol + 1).reshape(nc, ny, nx)
        out.append((box, a))
    return hdr, out


def gather(parts, shape, ncomp=1):
    """global array (ncomp, *shape) from the boxes of read_multifab"""
    g = np.full((ncomp,) + tuple(shape), np.nan)
    for (x0, y0, x1, y1, _, _), a in parts:
        g[:, y0:y1 + 1, x0:x1 + 1] = a[:ncomp]
    return g


def read_checkpoint_header(path: str):
    text = open(os.path.join(path, "Header")).read()
    lines = text.split("\n")
    return {"title": lines[0], "time": float(lines[1]), "boxes": parse_boxes(text)}


def read_plotfile_header(path: str):
    lines = open(os.path.join(path, "Header")).read().split("\n")
    nv = int(lines[1])
    names = lines[2:2 + nv]
    k = 2 + nv
    h = {"version": lines[0], "names": names, "dim": int(lines[k]), "time": float(lines[k + 1]),
         "finest_level": int(lines[k + 2]), "prob_lo": [float(x) for x in lines[k + 3].split()],
         "prob_hi": [float(x) for x in lines[k + 4].split()], "domain": lines[k + 6].strip(),
         "step": int(lines[k + 7]), "dx": [float(x) for x in lines[k + 8].split()], "coord": int(lines[k + 9])}
    lev = lines[k + 11].split()
    h["nboxes"] = int(lev[1])
    h["mf_path"] = lines[k + 13 + 2 * h["nboxes"]].strip()
    h["box_extents"] = [[float(x) for x in lines[k + 13 + r].split()] for r in range(2 * h["nboxes"])]
    return h
