"""Result tables in the reference's "FEMGPU Output File" format
(gpufem_amd/results.py, after src/Utils/fileIO.cc:408-600 and
src/interface.cpp:195-300): the reader parses two of the reference's own
result files (tests/golden/results/, copied by
tests/golden/nf/fetch_fixtures.sh), and a written file reads back to the
same numbers with the reference's layout, byte for byte where the layout
is fixed."""
import os

import pytest

from gpufem_amd import results as Rz

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "results")


def test_reads_reference_assembly_table():
    blocks = Rz.read_results(os.path.join(HERE, "Q4_col.txt"))
    b0 = blocks[0]
    assert b0["format"] == "CSR" and b0["gauss"] == 2 and b0["coloring"] == 1
    assert b0["threads"] == 1 and tuple(b0["columns"]) == Rz.ASM_COLUMNS
    assert b0["rows"][0] == [578, 289, 256, 0.073, 0.003, 0.007]
    assert b0["rows"][-1] == [132098, 66049, 65536, 18.047, 0.700, 0.306]
    assert [b["threads"] for b in blocks[:4]] == [1, 2, 3, 4]
    # the appended run after the '#' separator switches format
    assert blocks[4]["format"] == "ELL" and blocks[4]["rows"][0][3] == 0.141


def test_reads_reference_solve_table_without_trailing_tabs():
    blocks = Rz.read_results(os.path.join(HERE, "brk8_solve_col.txt"))
    assert tuple(blocks[0]["columns"]) == Rz.ASM_COLUMNS
    sol = blocks[1]
    assert tuple(sol["columns"]) == Rz.SOL_COLUMNS
    assert sol["rows"][0] == [14739, 4096, 10.795, 0.895, 0.330, 5.680, 1.721, 1.699, 1.872,
                              1.541, 1.984]


def test_written_file_layout_and_round_trip(tmp_path):
    out = str(tmp_path / "r.txt")
    with Rz.ResultsFile(out, gauss=3, coloring=True, date="Mon Jan  1 00:00:00 2024") as f:
        f.block(8, solve=False)
        f.row([578, 289, 256, 0.0731, 0.0025, 0.007])
        f.end_block()
    with Rz.ResultsFile(out, append=True) as f:
        f.block(1, solve=True)
        f.row([14739, 4096, 10.795, -1.0, 0.33, -1.0, 1.7, 1.69, 1.87, 1.54, 1.98])
        f.end_block()
    text = open(out).read()
    assert text.startswith("FEMGPU Output File\nDate: Mon Jan  1 00:00:00 2024\n"
                           "Gps:3\tColoring:1\t\nMatrix Format: ELL\nThreads:8\t\n"
                           "DOF\tNODES\tELEM\tSIZE\tAsmCPU\tAsmGPU\t\n"
                           "578\t289\t256\t0.073\t0.003\t0.007\t\n\n" + "#" * 59 + "\n")
    blocks = Rz.read_results(out)
    assert blocks[0]["gauss"] == 3 and blocks[0]["threads"] == 8
    assert blocks[1]["rows"][0] == [14739, 4096, 10.795, -1.0, 0.33, -1.0, 1.7, 1.69, 1.87,
                                    1.54, 1.98]


@pytest.mark.parametrize("dim", [2, 3])
def test_strategy_columns_map_to_fitting_variants(dim):
    v = Rz.STRATEGY_VARIANTS[dim]
    assert list(v) == list(Rz.SOL_COLUMNS[6:])
    # node-block and block-sliced variants exist per dimension (5/7: 2 DOF, 6/8: 3 DOF)
    assert v["SolBlk"] == (5 if dim == 2 else 6) and v["SolBlkUR"] == (7 if dim == 2 else 8)
