"""Writes the output-format fixtures of the Coacervate example from the files the
reference ships in examples/Coacervate (run in a container that holds the reference):

  coacervate/grid_densities.bin   the reference's file as shipped (its 32-byte header)
  coacervate/positions.bin.gz     the reference's file, gzip-compressed (header only:
                                  ns, Dim, L, types, molecules, charge flag, charges)
  coacervate/traj_step0.json      the step-0 frame of traj.lammpstrj as the SHA-256 of
                                  its whitespace-normalised lines, its line count and its
                                  first lines (the shipped file has CRLF line ends and one
                                  space before the charge column)
"""
import gzip
import hashlib
import json
import pathlib
import sys

SRC = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples/Coacervate")
OUT = pathlib.Path(__file__).resolve().parent / "coacervate"


def normalised_lines(raw):
    return [b" ".join(line.split()) for line in raw.replace(b"\r\n", b"\n").split(b"\n") if line.strip()]


def main():
    (OUT / "grid_densities.bin").write_bytes((SRC / "grid_densities.bin").read_bytes())
    with gzip.GzipFile(OUT / "positions.bin.gz", "wb", mtime=0) as f:
        f.write((SRC / "positions.bin").read_bytes())
    lines = normalised_lines((SRC / "traj.lammpstrj").read_bytes())
    digest = hashlib.sha256(b"\n".join(lines)).hexdigest()
    (OUT / "traj_step0.json").write_text(json.dumps(
        {"sha256_normalised": digest, "lines": len(lines), "head": [l.decode() for l in lines[:12]]}, indent=1) + "\n")


if __name__ == "__main__":
    main()
