#!/bin/sh
# Copies the reference's small neutral-file meshes (data files) used as
# fixtures by tests/test_nf.py.  Run from the repository root in a container
# that has the reference checkout at /root/reference.
set -e
SRC=/root/reference/test_models
DST=tests/golden/nf
for f in placa2_Q4.nf placa_Q4.nf 2Q8.nf placa2_Q8.nf brick_brk20.nf brick_brk20v2.nf brick_brk8v2.nf; do
  cp "$SRC/$f" "$DST/$f"
done
cp "$SRC/Brk8/rubik10_brk8.nf" "$DST/rubik10_brk8.nf"
# two result tables of the reference (format fixtures for tests/test_results.py)
mkdir -p tests/golden/results
cp /root/reference/Report/test_models_home/Q4/_Q4_col.txt tests/golden/results/Q4_col.txt
cp "$SRC/Brk8/_brk8_solve_col.txt" tests/golden/results/brk8_solve_col.txt
