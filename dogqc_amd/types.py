"""Relational data types of the plan API.

Same member names and values as the reference enum (dogqc/types.py:3-10) so
that plan and schema files written against it evaluate unchanged.  INT64 is an
extension (no reference counterpart): a 64-bit signed integer column, needed
for payloads whose products/sums exceed 32 bits.  DOUBLE columns are stored as
8-byte IEEE doubles in mmdb files (the reference never declares a DOUBLE base
column; its FLOAT columns are 4-byte, cudaTranslator.py:148-150).
"""
from enum import Enum


class Type(Enum):
    INT = 1
    FLOAT = 2
    DOUBLE = 3
    CHAR = 4
    STRING = 5
    DATE = 6
    BOOLEAN = 7
    INT64 = 8


# bytes per element of a base-table column of this type (STRING: offsets only)
COLUMN_BYTES = {
    Type.INT: 4,
    Type.FLOAT: 4,
    Type.DOUBLE: 8,
    Type.CHAR: 1,
    Type.DATE: 4,
    Type.INT64: 8,
    Type.BOOLEAN: 1,
}

# numpy dtype names of base-table columns (STRING handled separately)
COLUMN_DTYPE = {
    Type.INT: "int32",
    Type.FLOAT: "float32",
    Type.DOUBLE: "float64",
    Type.CHAR: "int8",
    Type.DATE: "uint32",
    Type.INT64: "int64",
    Type.BOOLEAN: "int8",
}
