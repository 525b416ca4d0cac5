"""The repository's top-level scripts parse (bench.py and the build/smoke entry
are run by the driver on the GPU machine without a prior import here)."""
import ast
import pathlib

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_top_level_scripts_parse():
    for name in ("bench.py", "__graft_entry__.py"):
        ast.parse((ROOT / name).read_text(), filename=name)
