#!/usr/bin/env python3
"""Generate the marching-cubes table data includes from the reference's text.

The triangle tables are data that the mesher must reproduce exactly: the
reference's `TriOffsets` (KinectFission/kifi/Mesher.cpp:239-261, 257 ints) and
`TriTable` (Mesher.cpp:264-475, 2,460 ints).  This script parses the numbers
out of that file (read as text; nothing is compiled or run), checks them
against their sha256 fingerprints (int32 little-endian), and writes

    oracle/mc_tables.inc                    -- for the CPU oracle
    merge_fusion_amd/csrc/mc_tables.inc     -- for the HIP mesher

Run it in a checkout that has the reference at /root/reference; the generated
files are committed, so nothing reads the reference at build or run time.
"""
import hashlib
import os
import re
import struct
import sys

REF = "/root/reference/KinectFission/kifi/Mesher.cpp"
SHA_OFFSETS = "928c98cc7322ed23d4ac768872d62c3d940d5feb07c4c4b6644597d469815420"
SHA_TABLE = "236f2ca201a686af008b4172209bddf548bb84000da239ceef9cc00a93758da6"


def fingerprint(vals):
    return hashlib.sha256(struct.pack("<%di" % len(vals), *vals)).hexdigest()


def parse_array(text, name):
    m = re.search(r"static int const %s\[\] = \{(.*?)\};" % name, text, re.S)
    if not m:
        raise SystemExit("array %s not found" % name)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [int(t) for t in re.findall(r"-?\d+", body)]


def emit(path, offsets, table, storage):
    lines = [
        "/* Generated by oracle/gen_mc_tables.py -- marching-cubes table data.",
        " * kTriOffsets: 257 entries, sha256(int32 LE) %s" % SHA_OFFSETS,
        " * kTriTable:  2460 entries, sha256(int32 LE) %s */" % SHA_TABLE,
        "%sconst int kTriOffsets[257] = {" % storage,
    ]
    for i in range(0, len(offsets), 16):
        lines.append("    " + ", ".join(str(v) for v in offsets[i:i + 16]) + ",")
    lines.append("};")
    lines.append("%sconst int kTriTable[2460] = {" % storage)
    for i in range(0, len(table), 24):
        lines.append("    " + ", ".join(str(v) for v in table[i:i + 24]) + ",")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    text = open(REF).read()
    offsets = parse_array(text, "triOffsets")
    table = parse_array(text, "triTable")
    assert len(offsets) == 257 and len(table) == 2460, (len(offsets), len(table))
    if fingerprint(offsets) != SHA_OFFSETS or fingerprint(table) != SHA_TABLE:
        raise SystemExit("fingerprint mismatch")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    emit(os.path.join(root, "oracle", "mc_tables.inc"), offsets, table, "static ")
    emit(os.path.join(root, "merge_fusion_amd", "csrc", "mc_tables.inc"), offsets, table,
         "static __constant__ ")
    print("ok: offsets %s table %s" % (fingerprint(offsets)[:12], fingerprint(table)[:12]))


if __name__ == "__main__":
    sys.exit(main())
