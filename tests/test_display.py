"""The record printer (gpufs_amd.display, the counterpart of the reference's
samples/utils/display.cpp) on hand-built records."""
import struct

from gpufs_amd import display


def _rec(word, count, name):
    return word.ljust(32, b"\0")[:32] + struct.pack("<i", count) + name.ljust(64, b"\0")


def test_lines_and_records_known_answers():
    data = _rec(b"hello", 1234, b"db1.txt") + _rec(b"_x", 1, b"/a/b,c") + _rec(b"w", -7, b"")
    assert display.lines(data) == [b"hello 1234 db1.txt\n", b"_x 1 /a/b,c\n", b"w -7 \n"]
    assert display.records(data) == [(b"hello", 1234, b"db1.txt"), (b"_x", 1, b"/a/b,c"), (b"w", -7, b"")]


def test_full_width_query_runs_into_the_count_like_printf():
    # display.cpp's first %s starts at the record and stops at the first NUL:
    # with 32 query bytes that is inside the count (5 = 05 00 00 00)
    data = _rec(b"A" * 32, 5, b"x")
    assert display.lines(data) == [b"A" * 32 + b"\x05 5 x\n"]
    assert display.records(data) == [(b"A" * 32, 5, b"x")]


def test_trailing_partial_record_is_ignored():
    data = _rec(b"a", 2, b"f") + b"\x01" * 40
    assert display.lines(data) == [b"a 2 f\n"]


def test_main_prints_usage_without_arguments(capsys):
    assert display.main([]) == 255
    assert "<input file" in capsys.readouterr().out
