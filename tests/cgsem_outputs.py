"""Readers for the output files of a CGSEM driver run (the reference's
CGsemElliptic_2D % WriteTecplot / WriteResidual / ConstructMatrix and
CGsemPreconditioner_2D % ConstructPCMatrix formats,
src/cgsem/CGsemElliptic_2D_Class.f90:1265-1453,
CGsemPreconditioner_2D_Class.f90:750-851): list-directed numbers, which the
compiler may wrap over several lines, so every file is read as one stream
of numbers after its header and zone lines."""
import os

import numpy as np


def _numbers(text):
    return np.array([float(t.replace("D", "E")) for t in text.split()])


def _tec(path, nvar):
    """(rows, nvar) values of a Tecplot POINT file; zone headers skipped."""
    lines = open(path).read().splitlines()[1:]
    body = " ".join(l for l in lines if "ZONE" not in l)
    v = _numbers(body)
    return v.reshape(-1, nvar)


def numeric_rows(path, nvar):
    """(rows, nvar) of every number in a list-directed text file, skipping
    the lines that are not all numbers (titles, zone headers, wrapped
    variable lists)."""
    vals = []
    for line in open(path).read().splitlines():
        toks = line.split()
        try:
            vals.extend(float(t.replace("D", "E")) for t in toks)
        except ValueError:
            continue
    return np.array(vals).reshape(-1, nvar)


def read_run(d):
    """Every output of one driver run in directory d, as arrays."""
    out = {}
    for name in ("Omat", "PCmat"):
        v = _numbers(open(os.path.join(d, f"{name}.txt")).read())
        n = int(round(np.sqrt(v.size)))
        assert n * n == v.size, (name, v.size)
        out[name] = v.reshape(n, n)
    out["omask"] = _tec(os.path.join(d, "omask.tec"), 3)
    out["mask"] = _tec(os.path.join(d, "mask.tec"), 3)
    out["sol"] = _tec(os.path.join(d, "CGsemSol.tec"), 4)
    rows = [l.split() for l in open(os.path.join(d, "residual.curve")).read().splitlines()[1:] if l.strip()]
    out["resi"] = np.array([float(r[1]) for r in rows])
    return out
