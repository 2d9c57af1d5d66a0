"""The reference-side half of the boundary: integration/HpgemAPIDevice.h is
the adaptor a maintainer adds to hpGEM. It must compile against the
reference's own kernel headers (HpgemAPILinear.h, HpgemAPISimplified.h,
ConfigurationData.h, Element.h, the output writers) with every member
function instantiated, for DIM = 2 and 3. Skipped where the reference tree
is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_KERNEL = "/root/reference/kernel"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_KERNEL) or shutil.which("g++") is None,
                                reason="needs the reference source tree and g++")


def test_adaptor_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "instantiate.cpp"
    src.write_text('#include "HpgemAPIDevice.h"\n'
                   "template class hpgem::Base::HpgemAPIDevice<2>;\n"
                   "template class hpgem::Base::HpgemAPIDevice<3>;\n")
    cmd = ["g++", "-std=c++14", "-fsyntax-only", "-Wall", "-Wextra", f"-I{REF_KERNEL}",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'integration')}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    # the adaptor itself adds no warnings (the reference's headers have their own)
    own = [ln for ln in r.stderr.splitlines() if "HpgemAPIDevice.h:" in ln and ": warning:" in ln]
    assert not own, "\n".join(own)


def test_adaptor_declares_only_entry_points_of_the_abi():
    """Every hgdg_* call in the adaptor is declared in include/hgdg.h."""
    import re

    hdr = open(os.path.join(ROOT, "include", "hgdg.h")).read()
    declared = set(re.findall(r"\b(hgdg_\w+)\s*\(", hdr))
    used = set(re.findall(r"\b(hgdg_\w+)\s*\(", open(os.path.join(ROOT, "integration", "HpgemAPIDevice.h")).read()))
    assert used and used <= declared, used - declared
