/*
 * dogqc_amd .tbl importer (libdqcimport.so): TPC-H/SSB dbgen text tables
 * into mmdb column files, on the host, multi-threaded.
 *
 * Replaces the reference's importer, a C++ program generated per database
 * and compiled at import time (dogqc/dbio.py:117-253, Importer.loadDatabase /
 * getCodeReadTable), with the row counts taken by `wc -l`
 * (dbio.py:124-138).  Parsing follows the reference's reader exactly:
 * csv.h's io::CSVReader<ncols + 1, trim_chars<' '>, no_quote_escape<'|'>>
 * with throw_on_overflow (dbio.py:94-99), parse_signed_integer /
 * parse_float / single-character parse (csv.h:921-1066), and the schema's
 * fixed-position toDate (query/schema/tpch.py:9-20).  FLOAT columns come out
 * bit-identical to the reference's (its float arithmetic step by step, not a
 * correctly rounded strtof); DOUBLE and INT64 columns are 8-byte (the
 * reference parses DOUBLE as float and has no INT64).
 *
 * Output files are the mmdb format the plan API reads (dogqc_amd/dbio.py;
 * mappedmalloc.h:11-82): an 8-byte little-endian header holding
 * 8 + payload + 1, the payload, one NUL.  A STRING column is two files,
 * <path>_offset (uint64 offs[rows + 1], offs[0] = 0) and <path>_char.
 */
#ifndef DOGQC_AMD_IMPORT_H
#define DOGQC_AMD_IMPORT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DQC_IMPORT_MAX_COLUMNS 64

/* column types (the mmdb element type in brackets) */
enum {
    DQC_COL_INT = 0,    /* int32   */
    DQC_COL_INT64 = 1,  /* int64   */
    DQC_COL_FLOAT = 2,  /* float   */
    DQC_COL_DOUBLE = 3, /* double  */
    DQC_COL_CHAR = 4,   /* int8    */
    DQC_COL_DATE = 5,   /* uint32 yyyymmdd */
    DQC_COL_STRING = 6  /* offsets + chars */
};

/* error kinds of dqc_import_report.error_kind */
enum {
    DQC_IMPORT_OK = 0,
    DQC_IMPORT_COLUMNS = 1,     /* a row without exactly ncols + 1 '|'-separated fields */
    DQC_IMPORT_DIGIT = 2,       /* a character that is not a digit in a number */
    DQC_IMPORT_OVERFLOW = 3,    /* an integer out of its type's range */
    DQC_IMPORT_SINGLE_CHAR = 4, /* a CHAR field that is not exactly one character */
    DQC_IMPORT_SHORT_DATE = 5,  /* a DATE field shorter than YYYY-MM-DD (the reference reads past it) */
    DQC_IMPORT_IO = 6           /* a file that cannot be opened, mapped or written */
};

typedef struct dqc_import_report {
    int64_t rows;                                /* rows imported */
    int64_t error_line;                          /* 1-based line of the first error, 0 if none */
    int32_t error_column;                        /* 0-based field of that error, -1 for COLUMNS/IO */
    int32_t error_kind;                          /* DQC_IMPORT_* */
    int64_t char_bytes[DQC_IMPORT_MAX_COLUMNS];  /* heap bytes of each STRING column */
    char message[256];
} dqc_import_report;

/* Rows of a .tbl file: its lines, counting a last line without '\n'
 * (the reference's `wc -l`, dbio.py:124-138, misses that one). */
int dqc_tbl_rows(const char* tbl_path, int64_t* rows);

/* Import `tbl_path` into ncols column files: column k (type types[k]) goes to
 * column_paths[k] (STRING: column_paths[k] + "_offset" / "_char").  Uses
 * `threads` threads (<= 0: one per hardware thread).  Returns 0, or a
 * negative DQC_E* code (include/dogqc_amd.h) with `report` saying why; on an
 * error no column file is left behind.  The first error in line order is the
 * one reported, as the reference's sequential reader would. */
int dqc_import_tbl(const char* tbl_path, int32_t ncols, const int32_t* types, const char* const* column_paths,
                   int32_t threads, dqc_import_report* report);

#ifdef __cplusplus
}
#endif
#endif
