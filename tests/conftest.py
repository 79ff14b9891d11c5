import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle", "_build"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle module (test infrastructure)."""
    import onionpir_oracle
    return onionpir_oracle


@pytest.fixture(scope="session")
def comp_params(O):
    return O.PirParams("CONFIG_N2048_K1_COMP")


def scheme_from_oracle(p):
    """opir_scheme fields for a composite oracle PirParams."""
    return dict(poly_degree=p.N, composite=1, q1=p.q1, q2=p.q2, ntt_root=p.w_crt, plain_mod=p.plain_mod,
                small_q=p.small_q, l_ep=p.l_ep, base_log2=p.base_log2, l_key=p.l_key,
                base_log2_key=p.base_log2_key, l_ks=p.l_ks, bv_base_log2=p.bv_base_log2)
