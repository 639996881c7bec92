"""Every repository path the design notes, README and integration guide cite
(under rift_amd/, tests/, oracle/, include/ or profiles/) exists."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH_RE = re.compile(r"\b((?:rift_amd|tests|oracle|include|profiles)/[\w./-]+"
                     r"\.(?:py|hip|hpp|h|cpp|c|md|json|txt))\b")


@pytest.mark.parametrize("doc", ["DESIGN.md", "README.md", "INTEGRATION.md"])
def test_cited_paths_exist(doc):
    text = open(os.path.join(ROOT, doc)).read()
    cited = sorted(set(PATH_RE.findall(text)))
    missing = [p for p in cited if not os.path.exists(os.path.join(ROOT, p))]
    assert not missing, missing
