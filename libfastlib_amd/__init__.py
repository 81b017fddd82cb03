"""libfastlib_amd -- MI355X-native FastBit query hot path.

WAH bitvector logic, relic/bin range evaluation and index build as gfx950
HIP kernels behind a C ABI (``include/fastbit_hip.h``,
``lib/libfastbit_hip.so``).  This package is the Python host binding of that
ABI; the C++ host binding is ``include/fastbit_hip.hpp``.
"""
from . import native  # noqa: F401
from .native import (AND, OR, XOR, MINUS, Bitvector, FastbitError, Index, Conjunctions, Expressions, RangeBatch,
                     EXPR_AND, EXPR_OR, EXPR_XOR, EXPR_NAND, EXPR_NOR, EXPR_NOT, EXPR_MINUS, bench_op,
                     bench_op_rot, bench_cnt_rot, lib, PLAN_AUTO, PLAN_LITERAL, PLAN_COMPRESSED, PLAN_SCAN,
                     ENC_FUZZ, ENC_ZONA, TL_EMPTY, TL_FINE, TL_ALL, TL_EQUAL, TL_SAME_BIN)  # noqa: F401

__all__ = ["AND", "OR", "XOR", "MINUS", "Bitvector", "FastbitError", "Index", "Conjunctions", "Expressions", "RangeBatch",
           "EXPR_AND", "EXPR_OR", "EXPR_XOR", "EXPR_NAND", "EXPR_NOR", "EXPR_NOT", "EXPR_MINUS", "bench_op",
           "bench_op_rot", "bench_cnt_rot", "lib", "PLAN_AUTO", "PLAN_LITERAL", "PLAN_COMPRESSED", "PLAN_SCAN",
           "ENC_FUZZ", "ENC_ZONA", "TL_EMPTY", "TL_FINE", "TL_ALL", "TL_EQUAL", "TL_SAME_BIN"]
