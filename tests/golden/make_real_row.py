"""Writes real_row_75ifg.json: the real-data interferogram row (complex phase
and perpendicular baselines, n_trial_wraps 0.725669801235199) that the
reference's test of the Python-path topofit uses as input
(tests/test_stage2_ported.py:376-537).  The numbers are read as data from that
file's two array literals; no reference code is imported or run.

    python tests/golden/make_real_row.py [/path/to/reference]
"""
import ast
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
SRC = REF / "tests" / "test_stage2_ported.py"
FUNC = "test_ps_topofit_single_matches_source_near_max_refinement_path"


def main():
    tree = ast.parse(SRC.read_text())
    fn = next(n for n in ast.walk(tree) if isinstance(n, ast.FunctionDef) and n.name == FUNC)
    arrays = {}
    for node in fn.body:
        if isinstance(node, ast.Assign) and isinstance(node.targets[0], ast.Name):
            name = node.targets[0].id
            if name in ("cpxphase", "bperp") and isinstance(node.value, ast.Call):
                arrays[name] = ast.literal_eval(node.value.args[0])
    cpx = arrays["cpxphase"]
    out = {
        "source": "tests/test_stage2_ported.py:376-537 (input row of the near-max refinement test)",
        "n_trial_wraps": 0.725669801235199,
        "cpx_re": [complex(v).real for v in cpx],
        "cpx_im": [complex(v).imag for v in cpx],
        "bperp": [float(v) for v in arrays["bperp"]],
    }
    dst = Path(__file__).with_name("real_row_75ifg.json")
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {dst} ({len(cpx)} interferograms)")


if __name__ == "__main__":
    main()
