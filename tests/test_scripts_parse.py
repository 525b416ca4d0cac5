"""The repository's top-level scripts parse (bench.py and the build/smoke entry
run on a GPU machine, where a syntax error would only show at run time)."""
import ast
import pathlib

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_top_level_scripts_parse():
    for name in ("bench.py", "__graft_entry__.py"):
        ast.parse((ROOT / name).read_text(), filename=name)
