"""Writes tests/golden/*.json: known-answer vectors that pin the oracle.

* philox_kat.json -- Philox4x32-10 known-answer vectors published with Random123
  (D. E. Shaw Research, file kat_vectors, "philox4x32 10" rows): counter, key, output.
  The reference's stateless generator follows that algorithm (device_fields.cuh:150-207).
* cell_sizing_kat.json -- the asserts of the reference's tests/test_base_particle.cu:52-91
  (two systems: boxes 10x5 and 20x5, cell grids 5x2 and 4x3).
Run: python tests/golden/make_golden.py
"""
import json
from pathlib import Path

HERE = Path(__file__).resolve().parent

PHILOX = [
    {"ctr": [0x00000000, 0x00000000, 0x00000000, 0x00000000], "key": [0x00000000, 0x00000000],
     "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
    {"ctr": [0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], "key": [0xffffffff, 0xffffffff],
     "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
    {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
     "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
]

CELL_SIZING = {
    "box_x": [10.0, 20.0], "box_y": [5.0, 5.0], "dim_x": [5, 4], "dim_y": [2, 3],
    "cell_size_x": [2.0, 5.0], "cell_size_y": [2.5, 5.0 / 3.0],
    "cell_inv_x": [1.0 / 2.0, 1.0 / 5.0], "cell_inv_y": [1.0 / 2.5, 3.0 / 5.0],
    "cell_system_start": [0, 10, 22], "n_cells": 22,
    "source": "tests/test_base_particle.cu:52-91",
}

if __name__ == "__main__":
    (HERE / "philox_kat.json").write_text(json.dumps(PHILOX, indent=1))
    (HERE / "cell_sizing_kat.json").write_text(json.dumps(CELL_SIZING, indent=1))
    print("wrote", HERE / "philox_kat.json", HERE / "cell_sizing_kat.json")
