"""Operator subpackage (mirrors commit.operator)."""
