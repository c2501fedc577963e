"""Mutation check of the oracle's build composition.  CPU only.

Each mutant is a copy of oracle/oracle.cpp with one plausible mistake in its
restatement of index.h:287-323, compiled into a temporary library; the
closed-form cases of test_oracle_build.py must catch every one of them:
  * append_lt:        `<` instead of `<=` at the append threshold (index.h:311)
  * reverse_no_add:   the reverse robustPrune without G[t] (add=true, index.h:70-75, 315)
  * rows_after_rev:   the batch's rows written after the reverse step instead
                      of before it (index.h:298-300)
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from conftest import ROOT
from test_oracle_build import CASES

SRC = os.path.join(ROOT, "oracle", "oracle.cpp")

UPDATE_BLOCK = """    for (uint64_t i = 0; i < B; i++) {  // update_neighbors (graph.h:72-81)
      uint32_t* row = G.row(inserts[i + lo]);
      if (new_out[i].size() > G.R) { err = -3; continue; }
      row[0] = (uint32_t)new_out[i].size();
      for (size_t j = 0; j < new_out[i].size(); j++) row[1 + j] = new_out[i][j];
    }
    if (err) return err;
"""
AFTER_REVERSE = "    if (gerr) return gerr;\n"


def _append_lt(s):
    old = "if (S.size() + row[0] <= R) {  // append_neighbors"
    assert s.count(old) == 1
    return s.replace(old, "if (S.size() + row[0] < R) {  // append_neighbors")


def _reverse_no_add(s):
    old = "robust_prune(P, G, t, cc, alpha, true, R, tie_mode, o);"
    assert s.count(old) == 1
    return s.replace(old, "robust_prune(P, G, t, cc, alpha, false, R, tie_mode, o);")


def _rows_after_reverse(s):
    assert s.count(UPDATE_BLOCK) == 1 and s.count(AFTER_REVERSE) == 1
    s = s.replace(UPDATE_BLOCK, "")
    return s.replace(AFTER_REVERSE, AFTER_REVERSE + UPDATE_BLOCK)


MUTANTS = {"append_lt": _append_lt, "reverse_no_add": _reverse_no_add, "rows_after_rev": _rows_after_reverse}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory, oracle):
    d = tmp_path_factory.mktemp("mutants")
    src = open(SRC).read()

    def make(name):
        cpp = d / f"{name}.cpp"
        so = d / f"lib{name}.so"
        cpp.write_text(MUTANTS[name](src))
        r = subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-ffp-contract=off", "-fPIC", "-shared",
                            "-o", str(so), str(cpp)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return name, str(so)

    with ThreadPoolExecutor(len(MUTANTS)) as ex:
        return dict(ex.map(make, MUTANTS))


def _failing_cases(oracle, path):
    failed = []
    with oracle.using(path):
        for case in CASES:
            try:
                case(oracle)
            except Exception:  # a wrong row or an error both count as caught
                failed.append(case.__name__)
    return failed


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_caught(oracle, mutant_libs, name):
    failed = _failing_cases(oracle, mutant_libs[name])
    print(f"mutant {name}: caught by {failed}")
    assert failed, f"mutant {name} passes every closed-form build case"
