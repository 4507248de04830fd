"""Printer for the grep_text sample's binary records (gpufs_grep_text_records),
the host-side counterpart of the reference's samples/utils/display.cpp.

A record is 100 bytes: the query record's 32 bytes as stored, the count as a
little-endian int32 at byte 32, the database name NUL-padded to 64 bytes at
byte 36.  display.cpp:23-25 prints each record with
`fprintf(stdout, "%s %d %s\\n", line, *(int*)&line[32], line + 36)`; the
first `%s` reads from the record's start to its first NUL, so a 32-byte
query (no NUL in its field) runs on into the count's bytes, and `lines()`
prints it the same way.
"""
import struct
import sys

RECORD_BYTES = 100


def records(data):
    """(word bytes, count, name bytes) for each whole 100-byte record of `data`."""
    out = []
    for off in range(0, len(data) - RECORD_BYTES + 1, RECORD_BYTES):
        r = data[off:off + RECORD_BYTES]
        word = r[:32].split(b"\0", 1)[0]
        count = struct.unpack_from("<i", r, 32)[0]
        name = r[36:].split(b"\0", 1)[0]
        out.append((word, count, name))
    return out


def lines(data):
    """The lines display.cpp prints for `data`, as bytes, in record order."""
    out = []
    for off in range(0, len(data) - RECORD_BYTES + 1, RECORD_BYTES):
        r = data[off:off + RECORD_BYTES]
        head = r.split(b"\0", 1)[0]   # %s from the record's start
        count = struct.unpack_from("<i", r, 32)[0]
        name = r[36:].split(b"\0", 1)[0]
        out.append(head + b" " + str(count).encode() + b" " + name + b"\n")
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("<input file 32 byte + 4 byte  >")
        return 255
    with open(argv[0], "rb") as f:
        data = f.read()
    sys.stdout.buffer.write(b"".join(lines(data)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
