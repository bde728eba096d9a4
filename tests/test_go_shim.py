"""CPU checks of the Go binding (go/mopsgpu), which cannot be compiled
here (no Go toolchain): every C function it calls is declared in
include/mops.h and exported by libmops.so, and every function of the
reference's Go API on this path has the same parameter and result types in
the shim as in the reference (mmult.go, chol.go, lu.go, ldlnp.go, trinv.go,
ldl.go, ldlbk.go, ldlbks.go;
parameter names may differ, e.g. C -> Cm since C names the cgo package).

The reference's signatures are kept here as data (a list of type strings),
read from its sources once: the GPU box has no copy of the reference.
"""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GO = os.path.join(ROOT, "go", "mopsgpu")

# name -> (parameter types, result types) of the reference's exported functions
REFERENCE_SIGNATURES = {
    "Mult": (["*matrix.FloatMatrix"] * 3 + ["float64"] * 2 + ["Flags"], ["error"]),
    "MultSym": (["*matrix.FloatMatrix"] * 3 + ["float64"] * 2 + ["Flags"], ["error"]),
    "MultTrm": (["*matrix.FloatMatrix"] * 2 + ["float64", "Flags"], ["error"]),
    "SolveTrm": (["*matrix.FloatMatrix"] * 2 + ["float64", "Flags"], ["error"]),
    "RankUpdateSym": (["*matrix.FloatMatrix"] * 2 + ["float64"] * 2 + ["Flags"], ["error"]),
    "RankUpdate2Sym": (["*matrix.FloatMatrix"] * 3 + ["float64"] * 2 + ["Flags"], ["error"]),
    "UpdateTrm": (["*matrix.FloatMatrix"] * 3 + ["float64"] * 2 + ["Flags"], ["error"]),
    "DecomposeCHOL": (["*matrix.FloatMatrix", "Flags", "int"], ["*matrix.FloatMatrix", "error"]),
    "SolveCHOL": (["*matrix.FloatMatrix"] * 2 + ["Flags"], []),
    "DecomposeLU": (["*matrix.FloatMatrix", "[]int", "int"], ["*matrix.FloatMatrix", "error"]),
    "SolveLU": (["*matrix.FloatMatrix"] * 2 + ["[]int", "Flags"], ["error"]),
    "DecomposeLUnoPiv": (["*matrix.FloatMatrix", "int"], ["*matrix.FloatMatrix", "error"]),
    "DecomposeLDLnoPiv": (["*matrix.FloatMatrix"] * 2 + ["Flags", "int"],
                          ["*matrix.FloatMatrix", "error"]),
    "SolveLDLnoPiv": (["*matrix.FloatMatrix"] * 2 + ["Flags"], []),
    "InverseTrm": (["*matrix.FloatMatrix", "Flags", "int"], ["*matrix.FloatMatrix", "error"]),
    "DecomposeLDL": (["*matrix.FloatMatrix"] * 2 + ["[]int", "Flags", "int"],
                     ["*matrix.FloatMatrix", "error"]),
    "SolveLDL": (["*matrix.FloatMatrix"] * 2 + ["[]int", "Flags"], []),
    "DecomposeBK": (["*matrix.FloatMatrix"] * 2 + ["[]int", "Flags", "int"],
                    ["*matrix.FloatMatrix", "error"]),
    "SolveBK": (["*matrix.FloatMatrix"] * 2 + ["[]int", "Flags"], []),
    "BlockingParams": (["int"] * 3, []),
    "NumWorkers": (["int"], ["int"]),
    "SetPanicOnError": (["bool"], []),
}


def _go_sources():
    return "\n".join(open(os.path.join(GO, f)).read() for f in sorted(os.listdir(GO))
                     if f.endswith(".go"))


def _param_types(params):
    """Types of a Go parameter list, expanding grouped names (a, b T)."""
    types, pending = [], 0
    for p in [p.strip() for p in params.split(",") if p.strip()]:
        parts = p.split()
        if len(parts) == 1:
            pending += 1
            continue
        types += [parts[-1]] * (pending + 1)
        pending = 0
    return types


def _signatures(src):
    out = {}
    for m in re.finditer(r"^func (\w+)\(([^)]*)\)\s*([^{]*)\{", src, re.M):
        res = m.group(3).strip()
        if res.startswith("("):
            res = [r.split()[-1] for r in res[1:-1].split(",") if r.strip()]
        else:
            res = [res] if res else []
        out[m.group(1)] = (_param_types(m.group(2)), res)
    return out


def test_shim_signatures_match_the_reference_api():
    sigs = _signatures(_go_sources())
    for name, want in REFERENCE_SIGNATURES.items():
        assert name in sigs, f"{name} missing from go/mopsgpu"
        assert sigs[name] == want, (name, sigs[name], want)


def test_shim_calls_only_declared_and_exported_symbols():
    from matops_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "mops.h")).read()
    declared = set(re.findall(r"\b(mops_\w+)\s*\(", hdr))
    used = set(re.findall(r"\bC\.(mops_\w+)\s*\(", _go_sources()))
    assert used, "no library calls found"
    assert used <= declared, sorted(used - declared)
    L = ctypes.CDLL(_lib.LIBPATH)
    assert all(hasattr(L, n) for n in used)


def test_shim_error_strings_are_the_reference_ones():
    src = _go_sources()
    for msg in ("Mult: size mismatch", "MultSym: size mismatch", "A, B size mismatch",
                "C not a square matrix", "A not a square matrix",
                "pivot array < min(A.Rows(),A.Cols())", "No workspace for blocking invocation",
                "Workspace too small"):
        assert f'"{msg}"' in src, msg
