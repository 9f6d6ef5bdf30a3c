"""bench.py read as source (no GPU): every function that main() calls is
defined at module level or imported, and every leg function exists."""
import ast
import builtins
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_main_calls_only_defined_functions():
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    defined = {n.name for n in tree.body if isinstance(n, ast.FunctionDef)}
    imported = set()
    for n in ast.walk(tree):
        if isinstance(n, (ast.Import, ast.ImportFrom)):
            imported |= {a.asname or a.name.split(".")[0] for a in n.names}
    main = next(n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name == "main")
    local = {t.id for n in ast.walk(main) if isinstance(n, (ast.Assign, ast.For))
             for t in ast.walk(n.targets[0] if isinstance(n, ast.Assign) else n.target) if isinstance(t, ast.Name)}
    called = {n.func.id for n in ast.walk(main) if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
    missing = called - defined - imported - local - set(dir(builtins))
    assert not missing, missing
    for leg in ("batched_run", "exact_run", "eagle_run", "lattice_run", "adjoint_timing", "plain_gpu_baseline",
                "cpu_baseline", "exact_cpu_baseline"):
        assert leg in defined, leg
