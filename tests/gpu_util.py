"""Helpers shared by the GPU tests."""
import gzip
import pathlib
import shutil

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def coacervate_dir(tmp):
    d = pathlib.Path(tmp) / "coacervate"
    d.mkdir(parents=True, exist_ok=True)
    shutil.copy(GOLDEN / "coacervate" / "input", d / "input")
    with gzip.open(GOLDEN / "coacervate" / "input.data.gz", "rb") as fi, open(d / "input.data", "wb") as fo:
        shutil.copyfileobj(fi, fo)
    return (d / "input").read_text(), d


def coacervate_golden():
    line = (GOLDEN / "coacervate" / "data.dat").read_text().splitlines()[1].split()
    v = [float(x) for x in line]
    return {"Upe": v[1], "Ubond": v[2], "Ptens": np.array(v[3:9]), "Ugauss0": v[9]}


def rel_norm(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def min_image(dx, L):
    L = np.asarray(L)
    return dx - L * np.round(dx / L)
