"""Writes tests/golden/data2D_025.txt.gz: the reference's own 2-D snapshot
config/data2D_025.txt (DataPrinter::print2D format, 150 x 200 cells),
gzip-compressed unchanged, for the snapshot format tests.

    python tests/golden/snapshot_fixture.py [/root/reference]
"""
import gzip
import os
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data2D_025.txt.gz")
with open(os.path.join(ref, "config", "data2D_025.txt"), "rb") as src, gzip.open(out, "wb", compresslevel=9) as dst:
    shutil.copyfileobj(src, dst)
print(out, os.path.getsize(out))
