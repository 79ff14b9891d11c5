"""Host API (C++ via pybind11): PirParams, PirClient, PirServer, GpuScheme, and the
continuous-batching ServingLoop over several resident databases.

Importing fails loudly if the native module has not been built
(`python -m onionpirv2_amd.build`).
"""
import os as _os
import sys as _sys

_LIB = _os.path.join(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))), "lib")
if _LIB not in _sys.path:
    _sys.path.insert(0, _LIB)
try:
    from _opir_host import (GpuScheme, PirClient, PirParams, PirServer, ServingLoop, chacha20_block,  # noqa: F401
                            decode_keys, decode_query, encode_keys, encode_query, expand_keys, gaussian_cdt,
                            keys_wire_bytes, pack_bits, plan_batch_limits, prg_gaussian, prg_ternary, prg_uniform,
                            query_seed, query_wire_bytes, unpack_bits, TreePirParams, TreePirServer,
                            make_tree_pir_params_for_scheme, tree_query_expansion_params, client_index,
                            build_level_plans, synthetic_tree_node_value, pack_tree_level_synthetic,
                            make_tree_query)
except ImportError as e:  # pragma: no cover
    raise ImportError(f"onionpirv2_amd native host module is not built in {_LIB}: {e}") from e
