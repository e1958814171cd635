"""The integration notes quote the snippet files that bin/integration_example
compiles (fvm-3d-f90_amd/fortran/integration/*.inc) verbatim, so the text a
maintainer pastes into the reference is the text that is tested on the GPU."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INTEG = os.path.join(ROOT, "fvm-3d-f90_amd", "fortran", "integration")
NAMES = ["getters_decl.inc", "getters.inc", "snippet_decl.inc", "snippet_setup.inc",
         "snippet_step.inc", "snippet_output.inc"]


def doc_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```fortran\n(.*?)```", text, flags=re.S)


@pytest.mark.parametrize("name", NAMES)
def test_snippet_quoted_verbatim(name):
    src = open(os.path.join(INTEG, name)).read()
    assert src in doc_blocks(), f"{name} is not quoted verbatim in INTEGRATION.md"


def test_example_includes_every_snippet():
    prog = open(os.path.join(INTEG, "integration_example.f90")).read()
    mod = open(os.path.join(INTEG, "ref_standin.f90")).read()
    for name in NAMES:
        assert f"include '{name}'" in prog + mod, name


def test_snippet_reads_no_private_member():
    """The snippet reaches the reference's types only through members that
    are public there: the scheme codes, gamma, block(:), idimension and the
    type-bound getters (kappa, limiter, entropy_fix_coeff and q are private,
    src/solve/class_space_discret.f90:29-32, src/base/defs_block_data.f90:
    28-36)."""
    text = "".join(open(os.path.join(INTEG, n)).read() for n in NAMES[2:])
    for private in ("%kappa", "%limiter", "%entropy_fix_coeff", "%q(", "%q ", "%q)"):
        assert private not in text, private
