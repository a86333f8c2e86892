"""The host program's front end: configuration file, CSV tokenizer, table
I/O (CPU, through a g++ build of the headers), and the psdbscan program end
to end against the oracle (GPU).

The configuration fixture tests/golden/psdbscan.conf is the reference's
data/psdbscan.conf; its values are read as boost's INFO parser documents
them (cpp/ConfigHolder.hpp:56-59).  The tokenizer cases follow the rules of
boost::escaped_list_separator (cpp/csvreader.hpp:47), the table behaviour
cpp/csvreader.hpp:101-172, cpp/TableHolder.hpp:146-212 and
cpp/csvwriter.hpp:37-73.
"""
from __future__ import annotations

import os
import random
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "immunodbscan_amd", "csrc")
CONF = os.path.join(ROOT, "tests", "golden", "psdbscan.conf")


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fe") / "frontend_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + CSRC, os.path.join(ROOT, "tests", "cpp", "frontend_check.cpp"),
                    "-o", exe], check=True)

    def run(*args, stdin=None):
        r = subprocess.run([exe, *map(str, args)], input=stdin, capture_output=True, text=True, timeout=60)
        return r.returncode, r.stdout

    return run


def test_reference_config_values(check):
    rc, out = check("conf", CONF, "input_settings.separator", "input_settings.endline", "input_settings.quotation",
                    "output_settings.quotation", "benchmark_output_settings.quotation", "gene_columns.first_gene.name",
                    "gene_columns.first_gene.ref", "gene_columns.second_gene.name", "sequence_columns.name",
                    "maximum_relative_edit_distance", "minimum_number_of_neighbors", "use_hamming",
                    "get_tree_depth_statistics", "sequences_per_kernel_block", "output_prefix", "output_suffix",
                    "output_filetype", "remove_noise_from_output", "benchmark_output_prefix", "vgenes_reference",
                    "verbose", "benchmark")
    assert rc == 0
    got = dict(ln.split("|", 1) for ln in out.splitlines())
    assert got == {
        "input_settings.separator": r"\x09", "input_settings.endline": r"\x0a", "input_settings.quotation": '"',
        "output_settings.quotation": r"\x00", "benchmark_output_settings.quotation": '"',
        "gene_columns.first_gene.name": "v_gene_column_name", "gene_columns.first_gene.ref": "data/VHgenes.txt",
        # second_gene and the sequence column are commented out in the file
        "gene_columns.second_gene.name": "<absent>", "sequence_columns.name": "<absent>",
        "maximum_relative_edit_distance": "0.15", "minimum_number_of_neighbors": "2", "use_hamming": "false",
        "get_tree_depth_statistics": "true", "sequences_per_kernel_block": "2", "output_prefix": "output_",
        "output_suffix": "", "output_filetype": ".csv", "remove_noise_from_output": "false",
        "benchmark_output_prefix": "benchmark_", "vgenes_reference": "VHgenes.txt", "verbose": "true",
        "benchmark": "true"}


def test_reference_config_column_selections_keep_file_order(check):
    rc, out = check("children", CONF, "select_columns_to_read")
    assert rc == 0
    assert [ln.split("|") for ln in out.splitlines()] == [
        ["nr_column_name", r"Sequence\x20number"], ["id_column_name", r"Sequence\x20ID"],
        ["v_gene_column_name", r"V-GENE\x20and\x20allele"], ["j_gene_column_name", r"J-GENE\x20and\x20allele"],
        ["d_gene_column_name", r"D-GENE\x20and\x20allele"], ["cdr3_column_name", "CDR3-IMGT"],
        ["vdj_column_name", "V-D-J-REGION"]]
    rc, out = check("children", CONF, "select_columns_to_write")
    assert [ln.split("|")[0] for ln in out.splitlines()] == [
        "file_origin", "nr_column_name", "id_column_name", "v_gene_column_name", "j_gene_column_name",
        "d_gene_column_name", "cdr3_column_name", "vdj_column_name", "cores", "copies", "clone_id"]
    assert out.splitlines()[-1] == "clone_id|Clone"


def test_config_includes_comments_quotes_and_braces(check, tmp_path):
    (tmp_path / "sub").mkdir()
    (tmp_path / "sub" / "extra.conf").write_text('inc_key "x y" ; from the included file\n')
    (tmp_path / "main.conf").write_text(
        '; comment line\n'
        '#include "sub/extra.conf"\n'
        'q "a\\tb;c\\"d"   ; the semicolon inside quotes is data\n'
        'bare value ; comment\n'
        'node\n{\n  k v\n  inner {\n    deep "1"\n  }\n}\n'
        'after 7\n')
    rc, out = check("conf", tmp_path / "main.conf", "inc_key", "q", "bare", "node.k", "node.inner.deep", "after",
                    "node.after")
    assert rc == 0
    assert out.splitlines() == ["inc_key|x\\x20y", 'q|a\\x09b;c"d', "bare|value", "node.k|v", "node.inner.deep|1",
                                "after|7", "node.after|<absent>"]
    (tmp_path / "bad.conf").write_text("node\n{\n k v\n")
    rc, out = check("conf", tmp_path / "bad.conf", "node.k")
    assert rc == 1 and "unmatched" in out


SPLIT_CASES = [
    # line, fields (None: the tokenizer rejects the line)
    ("a,b,c", ["a", "b", "c"]),
    ('a,"b,c",d', ["a", "b,c", "d"]),
    ('"a""b"', ["ab"]),                 # the quotation character toggles quoting and is dropped
    ("a\\,b,c", ["a,b", "c"]),          # escaped separator
    ('x\\"y', ['x"y']),                 # escaped quotation character
    ("a\\\\b", ["a\\b"]),               # escaped escape character
    ("a\\nb", ["a\nb"]),                # \n stands for a newline
    ("a,", ["a", ""]),                  # a final separator adds an empty field
    (",", ["", ""]),
    ("", []),                           # an empty line has no fields
    ("a\\xb", None),                    # unknown escape
    ("a\\", None),                      # escape at the end
]


def _show(s):
    return "".join(c if 0x20 < ord(c) < 0x7F and c not in "|\\" else f"\\x{ord(c):02x}" for c in s)


def test_split_fields_follows_escaped_list_rules(check):
    rc, out = check("split", ord(","), ord('"'), stdin="\n".join(c[0] for c in SPLIT_CASES) + "\n")
    assert rc == 0
    for (line, fields), got in zip(SPLIT_CASES, out.splitlines(), strict=True):
        want = "error" if fields is None else "|".join([str(len(fields))] + [_show(f) for f in fields])
        assert got == want, line
    rc, out = check("split", 9, ord('"'), stdin='id\t"s\tq"\tv\n')
    assert out.splitlines() == ["3|id|s\\x09q|v"]


def test_table_read_delete_write(check, tmp_path):
    src = tmp_path / "in.tsv"
    # a header with extra columns, a quoted field with the separator, CRLF
    # line ends, an empty sequence, and a final line break
    src.write_bytes(b'extra\tid\t"seq"\tv\r\n'
                    b'x\t1\tAC"G\t"T\tIGHV1\r\n'
                    b'x\t2\t\tIGHV2\r\n'
                    b'x\t3\tTTT\tIGHV3\r\n')
    dst = tmp_path / "out.tsv"
    rc, out = check("table", src, dst, 9, ord('"'), ord(","), ord("'"), "s", "s=seq", "i=id", "g=v")
    assert rc == 0, out
    # the final line break gives a fourth, empty, row; it goes with the empty sequence
    assert out.strip() == "read 4 deleted 2"
    assert dst.read_bytes() == (b"origin,seq,id,v\n"
                                + f"'{src}','ACG\tT','1','IGHV1'\n'{src}','TTT','3','IGHV3'\n".encode())
    rc, out = check("table", src, dst, 9, ord('"'), 9, 0, "s", "s=seq", "i=nope")
    assert rc == 1 and "Required column nope not found" in out


def test_table_missing_input_file(check, tmp_path):
    rc, out = check("table", tmp_path / "none.tsv", tmp_path / "o", 9, 34, 9, 0, "s", "s=seq")
    assert rc == 1 and "Cannot read from file" in out


# ---------------------------------------------------------------- GPU ----

HEAD = ["Sequence number", "Sequence ID", "V-GENE and allele", "J-GENE and allele", "D-GENE and allele",
        "CDR3-IMGT", "V-D-J-REGION", "Unused column"]


def _write_inputs(tmp_path, seed, vh, jh):
    from tests.test_oracle import random_rows

    rng = random.Random(seed)
    files = []
    rows = []
    for name, n in (("first.csv", 700), ("second.csv", 500)):
        recs = random_rows(rng, n, b"ACDEFGHIKLMNPQRSTVWY", 8, 22, len(vh), len(jh), 0.2, families=60, edits=2)
        lines = ["\t".join(f'"{h}"' for h in HEAD)]
        for k, (s, v, j) in enumerate(recs):
            vcall = ", or ".join(f"Homsap {vh[g][:-1]}*01 F" for g in v)
            jcall = ", or ".join(f"Homsap {jh[g][:-1]}*01 F" for g in j)
            cdr3 = s.decode()
            r = rng.random()
            if r < 0.03:
                cdr3 = ""
            elif r < 0.05:
                vcall = ""
            elif r < 0.07:
                jcall = ""
            lines.append("\t".join([str(k + 1), f'"id {k}"', vcall, jcall, "IGHD1-1*01", cdr3, "ACGT" * 3, "z"]))
            rows.append((name, str(k + 1), vcall, jcall, cdr3))
        (tmp_path / name).write_text("\n".join(lines) + "\n")
        files.append(name)
    return files, rows


def _expected(rows, vh, jh, eps, min_pts, vcol=True, jcol=True, scol=True):
    from oracle import oracle as O

    keep = [r for r in rows if (not scol or r[4]) and (not vcol or r[2]) and (not jcol or r[3])]
    vb = O.gene_bitsets([r[2].encode() for r in keep], [g.encode() for g in vh]) if vcol else \
        O.gene_bitsets([b"V"] * len(keep), [b"V"])
    jb = O.gene_bitsets([r[3].encode() for r in keep], [g.encode() for g in jh]) if jcol else \
        O.gene_bitsets([b"V"] * len(keep), [b"V"])
    seqs = [r[4].encode() if scol else b"Gene" for r in keep]
    lens = np.array([len(s) for s in seqs], np.uint32)
    off = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.uint64)
    exp = O.dbscan(O.Data(np.frombuffer(b"".join(seqs), np.uint8), off, lens, vb, jb), O.params(eps, -1, min_pts, False))
    return keep, exp


def _enable(text, *entries):
    """Uncomments the lines of the configuration whose text is one of entries."""
    out = []
    for ln in text.split("\n"):
        if ln.startswith(";") and ln[1:].strip() in entries:
            ln = ln[1:]
        out.append(ln)
    return "\n".join(out)


@pytest.mark.gpu
@pytest.mark.parametrize("remove_noise", [False, True])
def test_psdbscan_end_to_end(gpu_ctx, tmp_path, remove_noise):
    """psdbscan over two CSV files with the reference's configuration, the
    second gene and the CDR3 column switched on: rows with empty cells go,
    labels match the oracle, output columns, names and run log follow the
    reference program."""
    from immunodbscan_amd import _lib

    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "psdbscan")
    assert os.path.exists(exe), "psdbscan not built"
    vh = [ln for ln in open(os.path.join(ROOT, "tests/golden/VHgenes.txt")).read().split("\n") if ln]
    jh = [ln for ln in open(os.path.join(ROOT, "tests/golden/JHgenes.txt")).read().split("\n") if ln]
    (tmp_path / "data").mkdir()
    for g in ("VHgenes.txt", "JHgenes.txt"):
        (tmp_path / "data" / g).write_bytes(open(os.path.join(ROOT, "tests/golden", g), "rb").read())
    files, rows = _write_inputs(tmp_path, 31, vh, jh)
    conf = _enable(open(CONF).read(), "second_gene", "{", "}", "name\tj_gene_column_name",
                   'ref\t\t"data/JHgenes.txt"', "name  cdr3_column_name")
    conf = conf.replace("remove_noise_from_output \tfalse", f"remove_noise_from_output \t{str(remove_noise).lower()}")
    (tmp_path / "run.conf").write_text(conf)
    r = subprocess.run([exe, "--input", *files, "--config_file", "run.conf"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    keep, exp = _expected(rows, vh, jh, 0.15, 2)
    out = (tmp_path / "output_first.csv").read_text().split("\n")
    assert out[0] == "\t".join(["Original.Filename", "ID.number", "ID.code", "V_CALL", "J_CALL", "D_CALL", "CDR3",
                                "Full.Sequence", "Cores", "Copies", "Clone"])
    assert out[-1] == ""
    body = [ln.split("\t") for ln in out[1:-1]]
    clone = exp["clone_id"]
    sel = np.ones(len(keep), bool) if not remove_noise else clone != 0
    assert len(body) == int(sel.sum())
    kept = [k for k, s in zip(keep, sel) if s]
    assert [b[0] for b in body] == [k[0] for k in kept]
    assert [b[1] for b in body] == [k[1] for k in kept]
    assert [b[3] for b in body] == [k[2] for k in kept] and [b[6] for b in body] == [k[4] for k in kept]
    assert all(b[5] == "IGHD1-1*01" and b[7] == "ACGT" * 3 for b in body)
    assert [int(b[10]) for b in body] == clone[sel].tolist()
    assert [int(b[9]) for b in body] == exp["copies"][sel].tolist()
    assert [int(b[8]) for b in body] == exp["core"][sel].tolist()
    # the run log (cpp/psdbscan.cpp:80-103, cpp/TableHolder.hpp:189-300)
    log = (tmp_path / "benchmark_first.csv").read_text().split("\n")
    assert log[0] == "Timer/Message\tValue"
    cells = [ln.split("\t") for ln in log[1:] if ln]
    msgs = [b for a, b in cells if a.startswith("At ")]
    timers = [a for a, b in cells if not a.startswith("At ")]
    n_empty = [sum(1 for x in rows if not x[4]),
               sum(1 for x in rows if x[4] and not x[2]),
               sum(1 for x in rows if x[4] and x[2] and not x[3])]
    # the final line break of each file adds an empty row; the CDR3 pass deletes it
    want = ["Input file(s) first.csv", "Input file(s) second.csv", "Using configuration file run.conf",
            "Deleting empty rows for column cdr3_column_name", f"Deleting {n_empty[0] + 2} rows",
            "Deleting empty rows for column v_gene_column_name", f"Deleting {n_empty[1]} rows",
            "Deleting empty rows for column j_gene_column_name", f"Deleting {n_empty[2]} rows"]
    assert msgs[:9] == want
    assert f"Total number of sequences = {len(keep)}" in msgs
    assert f"Number of uniques = {exp['n_uniques']}" in msgs
    tail = msgs[msgs.index("Percentage of noise points = " + str(100 * int((clone == 0).sum()) // len(keep)) + "%"):]
    if remove_noise:
        assert tail[1:] == ["Deleting noise points", f"Deleting {int((clone == 0).sum())} rows",
                            "Writing to file output_first.csv"]
    else:
        assert tail[1:] == ["Writing to file output_first.csv"]
    assert timers[:3] == ["Reading Files", "Deleting Empty Rows", "Transferring data to device"]
    assert timers[-1] == "Overall"
    assert "Do clustering" not in timers and "Writing Output" not in timers


@pytest.mark.gpu
def test_psdbscan_reference_configuration(gpu_ctx, tmp_path):
    """The reference's configuration as shipped: V gene only, no sequence
    column (every row's sequence is "Gene"), output unquoted, the default
    data/psdbscan.conf and data/ExampleData.csv names."""
    from immunodbscan_amd import _lib

    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "psdbscan")
    vh = [ln for ln in open(os.path.join(ROOT, "tests/golden/VHgenes.txt")).read().split("\n") if ln]
    jh = [ln for ln in open(os.path.join(ROOT, "tests/golden/JHgenes.txt")).read().split("\n") if ln]
    (tmp_path / "data").mkdir()
    (tmp_path / "data" / "VHgenes.txt").write_bytes(open(os.path.join(ROOT, "tests/golden/VHgenes.txt"), "rb").read())
    (tmp_path / "data" / "psdbscan.conf").write_bytes(open(CONF, "rb").read())
    files, rows = _write_inputs(tmp_path, 32, vh, jh)
    os.replace(tmp_path / files[0], tmp_path / "data" / "ExampleData.csv")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [x for x in rows if x[0] == files[0]]
    keep, exp = _expected(rows, vh, jh, 0.15, 2, jcol=False, scol=False)
    out = (tmp_path / "data" / "output_ExampleData.csv").read_text().split("\n")
    body = [ln.split("\t") for ln in out[1:-1]]
    assert len(body) == len(keep)
    assert all(b[0] == "data/ExampleData.csv" for b in body)
    assert [int(b[10]) for b in body] == exp["clone_id"].tolist()
    assert [int(b[9]) for b in body] == exp["copies"].tolist()
    assert [int(b[8]) for b in body] == exp["core"].tolist()
    assert (tmp_path / "data" / "benchmark_ExampleData.csv").exists()
    r = subprocess.run([exe, "--help"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--config_file" in r.stdout
