"""numpy codecs of the matrix stream formats (same wire formats as the C++
codecs in csrc/jsvd_io.cpp and the reference's src/infrastructure/mat_file.cu,
txt_file.cu):

* .mat: records {u64 rows, u64 cols} + rows*cols doubles, all big-endian,
  row-major;
* .txt: whitespace-separated rows, matrices separated by blank lines.
"""
from __future__ import annotations

import struct

import numpy as np


def write_mat(path, mats) -> None:
    with open(path, "wb") as f:
        for a in mats:
            a = np.ascontiguousarray(a, dtype=np.float64)
            r, c = a.shape
            f.write(struct.pack(">QQ", r, c))
            f.write(a.astype(">f8").tobytes())


def read_mat(path):
    out = []
    with open(path, "rb") as f:
        while True:
            head = f.read(16)
            if not head:
                return out
            if len(head) != 16:
                raise ValueError("Truncated *.mat header.")
            r, c = struct.unpack(">QQ", head)
            body = f.read(8 * r * c)
            if len(body) != 8 * r * c:
                raise ValueError("Truncated *.mat payload.")
            out.append(np.frombuffer(body, dtype=">f8").astype(np.float64).reshape(r, c))


def write_txt(path, mats) -> None:
    with open(path, "w") as f:
        for k, a in enumerate(mats):
            if k:
                f.write("\n")
            for row in np.asarray(a, dtype=np.float64):
                f.write(" ".join(repr(float(x)) for x in row) + "\n")


def read_txt(path):
    out, cur = [], []
    with open(path) as f:
        for line in f:
            if not line.strip():
                if cur:
                    out.append(np.array(cur))
                    cur = []
                continue
            cur.append([float(x) for x in line.split()])
    if cur:
        out.append(np.array(cur))
    return out
