"""The ReaDDy-side adapter (readdy_amd/host/readdy_adapter/MI355XKernel.cpp) is compiled against an installed
ReaDDy, which this image lacks.  This CPU test checks it statically against the reference's headers when they
are present: every pure virtual member of Kernel, StateModel, ActionFactory and ObservableFactory is
overridden in the adapter, and every ABI function the adapter calls is declared in include/readdy_amd.h."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADAPTER = os.path.join(ROOT, "readdy_amd", "host", "readdy_adapter", "MI355XKernel.cpp")
REF = "/root/reference/include/readdy/model"
HEADERS = ["Kernel.h", "StateModel.h", "actions/ActionFactory.h", "observables/ObservableFactory.h"]


def _pure_virtuals(text):
    """names of the `virtual ... name(...) ... = 0;` declarations"""
    flat = re.sub(r"\s+", " ", re.sub(r"//[^\n]*|/\*.*?\*/", " ", text, flags=re.S))
    return set(re.findall(r"virtual [^;{}]*?\b(\w+)\s*\([^;{}]*\)[^;{}]*=\s*0\s*;", flat))


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference headers are not present")
def test_adapter_overrides_every_pure_virtual():
    src = open(ADAPTER).read()
    missing = {}
    for h in HEADERS:
        names = _pure_virtuals(open(os.path.join(REF, h)).read())
        assert names, h
        gone = sorted(n for n in names if not re.search(r"\b%s\s*\([^;{}]*\)[^;{}]*override" % n, src))
        if gone:
            missing[h] = gone
    assert not missing, missing


def test_adapter_calls_declared_abi_functions():
    src = open(ADAPTER).read()
    header = open(os.path.join(ROOT, "include", "readdy_amd.h")).read()
    called = set(re.findall(r"\b(rdy_\w+)\s*\(", src))
    declared = set(re.findall(r"\b(rdy_\w+)\s*\(", header))
    assert called and called <= declared, called - declared
