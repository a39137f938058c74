"""The drop-in commands (fastibs, fastibsmapper, KDBIntersect): CLI rules
and error paths on the CPU, byte-identical outputs on the GPU."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import corpus
from util import ROOT, sha256

import hipibs

BIN = os.path.join(ROOT, "gpu-ibs_amd", "bin")


def tool(name, *args, cwd=None):
    return subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True,
                          text=True, cwd=cwd, timeout=600)


@pytest.fixture(scope="module")
def small_layout(tmp_path_factory, small_corpus):
    """refs/small_genome.fasta and three dataset folders s0..s2."""
    d = tmp_path_factory.mktemp("layout")
    ref, samples = small_corpus
    (d / "refs").mkdir()
    (d / "refs" / "small_genome.fasta").write_bytes(
        corpus.fasta([(b"chr1", corpus.to_text(ref))]))
    (d / "refs" / "notes.txt").write_text("not a reference")
    for i, (keys, counts) in enumerate(samples):
        (d / f"s{i}").mkdir()
        hipibs.kmc_write(str(d / f"s{i}" / "sample"), keys, 25, counts=counts, p=9, sig_len=9,
                         counter_size=1, version=0x200, min_count=2, max_count=255)
    return d


def test_help_and_bad_arguments(tmp_path):
    for name in ("fastibs", "fastibsmapper"):
        r = tool(name, "--help")
        assert r.returncode == 0 and "Usage" in r.stdout
        assert tool(name).returncode == 1
    r = tool("fastibs", tmp_path, tmp_path, tmp_path, "50000abc")
    assert r.returncode == 1 and "Invalid <windowSize>" in r.stderr
    r = tool("fastibs", tmp_path / "nodb", tmp_path, tmp_path, "50000")
    assert r.returncode == 1 and "KMC database folder not found" in r.stderr
    r = tool("fastibs", tmp_path, tmp_path, tmp_path, "50000", "--cpu")
    assert r.returncode == 1 and "no CPU path" in r.stderr
    r = tool("fastibs", tmp_path, tmp_path, tmp_path, "50000", "--bogus")
    assert r.returncode == 1 and "unknown" in r.stderr
    for bad in ("--minimizer=4", "--minimizer=32", "--minimizer=21,0", "--minimizer=21,",
                "--minimizer=x"):
        r = tool("fastibs", tmp_path, tmp_path, tmp_path, "50000", bad)
        assert r.returncode == 1 and "--minimizer=M[,CAP]" in r.stderr, bad


def test_window_not_larger_than_k(small_layout, tmp_path):
    r = tool("fastibs", small_layout / "s0", small_layout / "refs", tmp_path / "out", 25)
    assert r.returncode == 1 and "must be larger than" in r.stderr


def test_no_gpu_is_an_error_not_a_fallback(small_layout, tmp_path):
    n, _ = hipibs.device_count()
    if n > 0:
        pytest.skip("a GPU is present")
    r = tool("fastibs", small_layout / "s0", small_layout / "refs", tmp_path / "out", 50000,
             cwd=tmp_path)
    assert r.returncode == 1
    assert "no usable GPU" in (tmp_path / "log.txt").read_text()
    assert not any(p.suffix == ".tsv" for p in (tmp_path / "out").iterdir())


@pytest.mark.gpu
def test_fastibs_matches_golden_and_skips_existing(small_layout, tmp_path):
    out = tmp_path / "out"
    r = tool("fastibs", small_layout / "s1", small_layout / "refs", out, 50000, "--batch-mb=1")
    assert r.returncode == 0, r.stderr
    got = (out / "s1_v_small_50000.tsv").read_bytes()
    assert got == (
        open(os.path.join(ROOT, "tests", "golden", "small_s1_50000.tsv"), "rb").read())
    r = tool("fastibs", small_layout / "s1", small_layout / "refs", out, 50000)
    assert r.returncode == 0 and "already exists" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("resident,opts", [(2, ()), (1, ()), (2, ("--minimizer=21",)),
                                           (1, ("--minimizer=13,4", "--load-factor=0.3"))])
def test_fastibs_panel_equals_singles(small_layout, tmp_path, resident, opts):
    """Panel runs in 2 or 3 memory groups (the next group's tables built
    while the current one scans) give the single-sample TSVs, over
    whole-k-mer tables and over minimizer tables (the panel kernel's
    grouped probe)."""
    out = tmp_path / "panel"
    lst = tmp_path / "list.txt"
    lst.write_text(f"# panel\n{small_layout / 's2'}\n")
    r = tool("fastibs", small_layout / "s0", small_layout / "refs", out, 50000,
             f"--db={small_layout / 's1'}", f"--db-list={lst}", f"--resident={resident}",
             *opts)
    assert r.returncode == 0, r.stderr
    for i in range(3):
        got = (out / f"s{i}_v_small_50000.tsv").read_bytes()
        want = open(os.path.join(ROOT, "tests", "golden", f"small_s{i}_50000.tsv"), "rb").read()
        assert got == want, i


@pytest.mark.gpu
def test_fastibsmapper_matches_golden(small_layout, tmp_path, expected):
    out = tmp_path / "map"
    r = tool("fastibsmapper", small_layout / "s0", small_layout / "refs", out)
    assert r.returncode == 0, r.stderr
    assert sha256((out / "s0_v_small_genome.txt").read_bytes()) == \
        expected["small"]["s0"]["coverage_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("opt", ["--minimizer=21", "--minimizer=13,4"])
def test_tools_with_minimizer_tables_match_golden(small_layout, tmp_path, expected, opt):
    """fastibs and fastibsmapper over minimizer tables (the grouped scan and
    hit-mask kernels) write the whole-k-mer tables' bytes."""
    out = tmp_path / "out"
    r = tool("fastibs", small_layout / "s1", small_layout / "refs", out, 50000, opt)
    assert r.returncode == 0, r.stderr
    assert (out / "s1_v_small_50000.tsv").read_bytes() == (
        open(os.path.join(ROOT, "tests", "golden", "small_s1_50000.tsv"), "rb").read())
    r = tool("fastibsmapper", small_layout / "s0", small_layout / "refs", out, opt)
    assert r.returncode == 0, r.stderr
    assert sha256((out / "s0_v_small_genome.txt").read_bytes()) == \
        expected["small"]["s0"]["coverage_sha256"]


@pytest.mark.gpu
def test_kdbintersect_matches_golden(small_layout, expected):
    for pair, want in expected["small"]["intersect"].items():
        a, b = pair.split("-")
        r = tool("KDBIntersect", small_layout / f"s{a}", small_layout / f"s{b}")
        assert r.returncode == 0, r.stderr
        assert r.stdout == f"{want}\n"


@pytest.mark.gpu
def test_tools_with_k_above_32(tmp_path, expected):
    """k = 47 databases (two-word keys): the fastibs TSV and the
    fastibsmapper file equal the reference's string path; a panel run equals
    single runs; KDBIntersect counts the shared keys."""
    k = 47
    recs, keys = corpus.edge_config(k)
    (tmp_path / "refs").mkdir()
    (tmp_path / "refs" / "edge_genome.fasta").write_bytes(corpus.fasta(recs))
    for name, ks in (("w0", keys), ("w1", keys[::2])):
        (tmp_path / name).mkdir()
        hipibs.kmc_write(str(tmp_path / name / "sample"), ks, k, counts=None, sig_len=9,
                         counter_size=1, version=0x200, min_count=1, max_count=255)
    e = expected["edge_wide"][str(k)]
    r = tool("fastibs", tmp_path / "w0", tmp_path / "refs", tmp_path / "one", 1000)
    assert r.returncode == 0, r.stderr
    single0 = (tmp_path / "one" / "w0_v_edge_1000.tsv").read_bytes()
    assert sha256(single0) == e["tsv"]["1000"]["sha256"]
    r = tool("fastibs", tmp_path / "w1", tmp_path / "refs", tmp_path / "one", 1000)
    assert r.returncode == 0, r.stderr
    r = tool("fastibs", tmp_path / "w0", tmp_path / "refs", tmp_path / "panel", 1000,
             f"--db={tmp_path / 'w1'}")
    assert r.returncode == 0, r.stderr
    for name in ("w0", "w1"):
        assert (tmp_path / "panel" / f"{name}_v_edge_1000.tsv").read_bytes() == \
            (tmp_path / "one" / f"{name}_v_edge_1000.tsv").read_bytes(), name
    r = tool("fastibsmapper", tmp_path / "w0", tmp_path / "refs", tmp_path / "map")
    assert r.returncode == 0, r.stderr
    assert sha256((tmp_path / "map" / "w0_v_edge_genome.txt").read_bytes()) == \
        e["coverage_sha256"]
    r = tool("KDBIntersect", tmp_path / "w0", tmp_path / "w1")
    assert r.returncode == 0, r.stderr
    assert r.stdout == f"{keys[::2].shape[0]}\n"


@pytest.mark.gpu
def test_kmerdatabase_members_match_golden(small_layout, tmp_path, expected):
    """The KmerDatabase members the tools do not call (per-contig mapping,
    chunked mapping into a difference array + serializeCoverage, the
    per-window multi-database and single-database statistics) give the
    golden coverage file and TSVs; DeviceKmerIndex's probeHost and
    probeCountHost give the golden intersect count."""
    out = tmp_path / "api"
    out.mkdir()
    dbs = [small_layout / f"s{i}" / "sample" for i in range(3)]
    r = tool("test_api", small_layout / "refs" / "small_genome.fasta", out, 50000, *dbs)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ("needsStringPath(25) 0, needsStringPath(64) 0, needsStringPath(65) 1, "
            "devicesToUse(4) 1, gpuRequested 1, "
            "gpuSyncRequested 0, poolThreads>0 1") in r.stdout
    n01 = expected["small"]["intersect"]["0-1"]
    assert f"intersect 0-1: probeCountHost {n01} probeHost {n01} devices 1" in r.stdout
    want_cov = expected["small"]["s0"]["coverage_sha256"]
    assert sha256((out / "mapping.txt").read_bytes()) == want_cov
    assert sha256((out / "mapping_chunks.txt").read_bytes()) == want_cov
    for i in range(3):
        want = open(os.path.join(ROOT, "tests", "golden", f"small_s{i}_50000.tsv"), "rb").read()
        assert (out / f"multi_{i}.tsv").read_bytes() == want, i
        assert (out / f"single_{i}.tsv").read_bytes() == want, i


def test_cpp_host_layer_checks():
    r = tool("test_host")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host checks ok" in r.stdout


def reader_records(path):
    """Records the k-mer counter's reader takes from a read file."""
    r = subprocess.run([os.path.join(BIN, "test_host"), "reads", str(path)], capture_output=True,
                       timeout=60)
    return r.returncode, r.stdout, r.stderr.decode()


def test_read_file_reader(tmp_path):
    """FASTQ with multi-line sequence and quality (quality lines starting with
    '+' and '@'), empty lines and an empty record; multi-line FASTA; gzip;
    CRLF kept as a non-base byte; bad input refused."""
    fq = (b"@r1\nACGT\n+\n@@@@\n\n@r2 x\nAC\ngt\n+r2\n+I\nII\n@r3\n+\n@r4\nNNTT\n+\n"
          b"@III\n")
    cases = {
        "a.fq": (fq, b"ACGT\nACgt\n\nNNTT\n", 4),
        "a.fq.gz": (gzip.compress(fq), b"ACGT\nACgt\n\nNNTT\n", 4),
        "b.fa": (b">a\nAC\nGT\n\n>b\n>c desc\nNNNN\r\nac\n", b"ACGT\n\nNNNN\rac\n", 3),
        "c.fa": (b">only\n" + b"ACGT" * 5000 + b"\n", b"ACGT" * 5000 + b"\n", 1),
    }
    for name, (data, want, n) in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        rc, out, err = reader_records(p)
        assert rc == 0, (name, err)
        assert out == want, name
        assert f"records {n}" in err, name
    for name, data, msg in (("d.txt", b"ACGT\n", "unknown file type"),
                            ("e.fq", b"@r\nAC\n+\nII\nGG\n", "malformed FASTQ"),
                            ("f.fq", b"", "empty file")):
        p = tmp_path / name
        p.write_bytes(data)
        rc, _, err = reader_records(p)
        assert rc == 2 and msg in err, (name, err)
    rc, _, err = reader_records(tmp_path / "missing.fq")
    assert rc == 2 and "cannot open" in err
