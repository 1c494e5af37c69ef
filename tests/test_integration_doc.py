"""INTEGRATION.md quotes the C++ calls of tests/integration/*.cpp verbatim
(the regions between `// [doc:NAME]` and `// [/doc:NAME]`), so the documented
binding is the code the GPU tests compile and run."""
import os
import re

from conftest import ROOT


def _regions(path):
    text = open(path).read()
    return {m.group(1): m.group(2) for m in re.finditer(r"// \[doc:(\w+)\]\n(.*?)\n\s*// \[/doc:\1\]", text, re.S)}


def _dedent(block):
    lines = block.splitlines()
    pad = min(len(l) - len(l.lstrip()) for l in lines if l.strip())
    return "\n".join(l[pad:] for l in lines)


def test_integration_doc_quotes_the_tested_programs():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\+\+\n(.*?)\n```", doc, re.S)
    found = {}
    for name in ("mc_host.cpp", "fdk_host.cpp"):
        found.update(_regions(os.path.join(ROOT, "tests", "integration", name)))
    assert set(found) == {"mc", "fdk"}
    for name, region in found.items():
        assert any(_dedent(region) in b for b in blocks), f"INTEGRATION.md does not quote the [doc:{name}] region"
