"""Regenerates the simple-point table entries of reference_pins.json
(sp_table_2d_bits, sp_table_3d_sha256, sp_table_3d_popcount) from the
oracle's definition-based tables.  The other entries of that file are
values copied from the reference repository's own tests and notebooks
(each entry names its source)."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

path = os.path.join(HERE, "reference_pins.json")
pins = json.load(open(path))
w = O.sp_table3d_words()
pins["sp_table_2d_bits"] = "".join(str(int(b)) for b in O.sp_table2d())
pins["sp_table_3d_sha256"] = hashlib.sha256(w.astype("<u8").tobytes()).hexdigest()
pins["sp_table_3d_popcount"] = int(sum(bin(int(x)).count("1") for x in w))
json.dump(pins, open(path, "w"), indent=1)
