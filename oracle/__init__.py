"""Checker side of the BC1 encoder: the CPU restatement (``icbc_oracle.c``)
and, where it was built, the reference encoder compiled from its own header
(``_ref/``).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py; the product (``icbc_amd``) never imports it.
"""
from .binding import Oracle, Reference, reference_available, build  # noqa: F401
