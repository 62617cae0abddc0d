#!/bin/sh
# Copies the reference's neutral-file meshes (data files) used as fixtures by
# tests/test_nf.py and tests/test_gpu_nf.py, and result tables of the
# reference (format fixtures for tests/test_results.py and the numbers the
# reference-mesh sweep prints beside its own).  Run from the repository root
# in a container that has the reference checkout at /root/reference.
set -e
SRC=/root/reference/test_models
DST=tests/golden/nf
for f in placa2_Q4.nf placa_Q4.nf 2Q8.nf placa2_Q8.nf brick_brk20.nf brick_brk20v2.nf brick_brk8v2.nf; do
  cp "$SRC/$f" "$DST/$f"
done
cp "$SRC/Brk8/rubik10_brk8.nf" "$DST/rubik10_brk8.nf"
# the meshes of the reference's published assembly tables
cp "$SRC/Brk8/rubik18_brk8.nf" "$DST/rubik18_brk8.nf"
cp "$SRC/Brk20/rubik20_brk20.nf" "$DST/rubik20_brk20.nf"
cp "$SRC/Q4/placa128_Q4.nf" "$DST/placa128_Q4.nf"
cp "$SRC/Q8/placa128_Q8.nf" "$DST/placa128_Q8.nf"
mkdir -p tests/golden/results
cp /root/reference/Report/test_models_home/Q4/_Q4_col.txt tests/golden/results/Q4_col.txt
cp "$SRC/Brk8/_brk8_solve_col.txt" tests/golden/results/brk8_solve_col.txt
cp /root/reference/Report/test_models_home/Q8/_Q8_col.txt tests/golden/results/Q8_col.txt
cp /root/reference/Report/test_models_home/Brk8/_brk8_col.txt tests/golden/results/brk8_col.txt
cp /root/reference/Report/test_models_home/Brk20/_brk20_col.txt tests/golden/results/brk20_col.txt
