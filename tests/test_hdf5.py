"""HDF5 read / write in the ITensors.jl / NDTensors layout (host side, no GPU).

Follows the reference's test/base/test_readwrite.jl: TagSet, Index (plain
and primed), IndexSet, Dense ITensor (real, compressed, complex), QN
ITensor (real, complex), and reading the v0.1.41 file (old "store" key).

Pins: the reference's own fixture (tests/golden/hdf5/testfilev0.1.41.h5,
its contents as read by h5dump in tests/golden/hdf5/testfilev0.1.41.json);
the header of a file this module writes, as dumped by h5dump, equals the
fixture's (same groups, datasets, attributes, datatypes and dataspaces);
h5dump reads the QN layout this module writes (variable-length names,
1-based block coordinates followed by 0-based offsets).
"""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import helpers  # noqa: F401  (puts the package on sys.path)

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "hdf5", "testfilev0.1.41.h5")
FIXTURE_JSON = os.path.join(HERE, "golden", "hdf5", "testfilev0.1.41.json")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module")
def env():
    import itg
    from itg import hdf5 as H5
    try:
        H5._h5()
    except H5.H5Error as e:
        pytest.skip(f"libhdf5 not available: {e}")
    return itg, H5


def ref_qn_inds(itg):
    """test_readwrite.jl:152-156."""
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(("A", -1)), 3), (QN(("A", 0)), 4), (QN(("A", 1)), 3)], "i")
    j = Index([(QN(("A", -2)), 2), (QN(("A", 0)), 3), (QN(("A", 2)), 2)], "j")
    k = Index([(QN(("A", -1)), 1), (QN(("A", 0)), 1), (QN(("A", 1)), 1)], "k")
    return i, j, k


def same_index(a, b):
    return (a == b and a.dir == b.dir and a.dim == b.dim
            and (a.space == b.space if not a.hasqns else
                 [(q, d) for q, d in a.space] == [(q, d) for q, d in b.space]))


def test_tag_order_follows_the_reference(env):
    """Tags sorted by their packed-character order, whitespace dropped,
    repeats dropped, at most four (src/lib/TagSets/src/TagSets.jl)."""
    itg, H5 = env
    from itg.itensors import tagstring
    assert tagstring("Site,S=1") == "S=1,Site"
    assert tagstring("A,Site,n=2") == "A,Site,n=2"
    assert tagstring(" b , a,a,,c,d,e") == "a,b,c,d"
    assert tagstring("Link,l=10") == "Link,l=10"
    assert tagstring("x,xy,X") == "X,x,xy"  # a prefix sorts first
    assert tagstring("abcdefghijklmnopqr") == "abcdefghijklmnop"  # 16 characters kept


def test_tagset_index_indexset_round_trip(env, tmp_path):
    """test_readwrite.jl:13-58."""
    itg, H5 = env
    Index = itg.Index
    p = str(tmp_path / "data.h5")
    ts = H5.TagSet("A,Site,n=2")
    i = Index(3, "Site,S=1")
    ip = i.prime(2)
    q = ref_qn_inds(itg)[0]
    two = Index([(itg.QN(("Sz", 1), ("Nf", 1)), 2), (itg.QN(("Nf", 0)), 1)], "two", dir=itg.In)
    iset = (Index(2, "i"), Index(3, "j"), Index(4, "k"))
    with H5.h5open(p, "w") as f:
        H5.write(f, "tags", ts)
        H5.write(f, "index", i)
        H5.write(f, "primed", ip)
        H5.write(f, "qn", q)
        H5.write(f, "two", two)
        H5.write(f, "inds", iset)
    with H5.h5open(p) as f:
        assert H5.read(f, "tags", H5.TagSet) == ts
        assert same_index(H5.read(f, "index", Index), i)
        rp = H5.read(f, "primed", Index)
        assert same_index(rp, ip) and rp.plev == 2
        assert same_index(H5.read(f, "qn", Index), q)
        rt = H5.read(f, "two", Index)
        assert same_index(rt, two) and rt.dir == itg.In
        ris = H5.read(f, "inds", H5.IndexSet)
        assert len(ris) == 3 and all(same_index(a, b) for a, b in zip(ris, iset))


@pytest.mark.parametrize("cplx,compress", [(False, None), (False, 3), (True, None)])
def test_dense_itensor_round_trip(env, tmp_path, cplx, compress):
    """test_readwrite.jl:62-106 (bit-exact here: the data is copied)."""
    itg, H5 = env
    Index = itg.Index
    T = itg.random_itensor(Index(2, "i"), Index(3, "j"), Index(4, "k"),
                           dtype=np.complex128 if cplx else np.float64,
                           rng=np.random.default_rng(1), device="cpu")
    p = str(tmp_path / "data.h5")
    with H5.h5open(p, "w") as f:
        H5.write(f, "T", T, compress=compress)
    with H5.h5open(p) as f:
        R = H5.read(f, "T", itg.ITensor, device="cpu")
    assert all(same_index(a, b) for a, b in zip(R.inds, T.inds))
    assert R.tensor.data.dtype == T.tensor.data.dtype
    assert np.array_equal(R.tensor.data.numpy(), T.tensor.data.numpy())


@pytest.mark.parametrize("cplx", [False, True])
def test_qn_itensor_round_trip(env, tmp_path, cplx):
    """test_readwrite.jl:152-180."""
    itg, H5 = env
    i, j, k = ref_qn_inds(itg)
    T = itg.random_itensor(itg.QN(("A", 1)) if not cplx else itg.QN(), i, j, k,
                           dtype=np.complex128 if cplx else np.float64,
                           rng=np.random.default_rng(2), device="cpu")
    p = str(tmp_path / "data.h5")
    with H5.h5open(p, "w") as f:
        H5.write(f, "T", T)
    with H5.h5open(p) as f:
        R = H5.read(f, "T", itg.ITensor, device="cpu")
    assert all(same_index(a, b) for a, b in zip(R.inds, T.inds))
    assert list(R.tensor.blockoffsets.items()) == list(T.tensor.blockoffsets.items())
    assert R.tensor.blockdims == T.tensor.blockdims
    assert np.array_equal(R.tensor.data.numpy(), T.tensor.data.numpy())
    assert R.flux() == T.flux()


@pytest.mark.parametrize("cplx", [False, True])
def test_delta_and_diag_itensor_round_trip(env, tmp_path, cplx):
    """test_readwrite.jl:108-150: delta (one value, Diag{T,T}) and a
    diagonal ITensor (Diag{T,Vector{T}}) without QNs."""
    itg, H5 = env
    k = itg.Index(4, "k")
    dt = np.complex128 if cplx else np.float64
    v = np.random.default_rng(5).standard_normal(4) + (1j if cplx else 0) * 0.5
    D = itg.diag_itensor(v.astype(dt), k, k.prime(), device="cpu")
    d = itg.delta(k, k.prime(), dtype=dt)
    p = str(tmp_path / "data.h5")
    with H5.h5open(p, "w") as f:
        H5.write(f, "diag_tensor", D)
        H5.write(f, "delta_tensor", d)
    with H5.h5open(p) as f:
        rD = H5.read(f, "diag_tensor", itg.ITensor, device="cpu")
        rd = H5.read(f, "delta_tensor", itg.ITensor, device="cpu")
        assert H5._attr(f.group("delta_tensor").group("storage"), "type") == (
            "Diag{ComplexF64,ComplexF64}" if cplx else "Diag{Float64,Float64}")
    assert rD.isdiag() and rd.isdiag() and rd.tensor.uniform == d.tensor.uniform
    assert np.array_equal(rD.array(), D.array()) and np.array_equal(rd.array(), d.array())
    assert all(same_index(a, b) for a, b in zip(rD.inds, D.inds))


def test_qn_diagonal_storage_is_refused(env, tmp_path):
    """No HDF5 form exists for DiagBlockSparse in the reference."""
    itg, H5 = env
    i = ref_qn_inds(itg)[0]
    with H5.h5open(str(tmp_path / "d.h5"), "w") as f:
        with pytest.raises(TypeError):
            H5.write(f, "d", itg.delta(i, i.prime().dag()))


def test_reads_the_reference_v0_1_41_file(env):
    """test_readwrite.jl:182-192: an ITensor written by ITensors.jl v0.1.41
    (storage under the old key "store") reads back with the ids, dims,
    tags and values h5dump shows."""
    itg, H5 = env
    exp = json.load(open(FIXTURE_JSON))
    with H5.h5open(FIXTURE) as f:
        assert "store" in f.group(exp["group"])
        T = H5.read(f, exp["group"], itg.ITensor, device="cpu")
    assert [i.id for i in T.inds] == [int(x["id"]) for x in exp["inds"]]
    assert [i.dim for i in T.inds] == [x["dim"] for x in exp["inds"]]
    assert [i.dir for i in T.inds] == [x["dir"] for x in exp["inds"]]
    assert [i.plev for i in T.inds] == [x["plev"] for x in exp["inds"]]
    assert [i.tags for i in T.inds] == [x["tags"] for x in exp["inds"]]
    assert np.array_equal(T.tensor.data.numpy(), np.array(exp["data"]))


def _header(path):
    out = subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True, text=True).stdout
    return out.split("\n", 1)[1]  # drop the line naming the file


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_written_layout_matches_the_reference_file(env, tmp_path):
    """The same tensor written by this module and by ITensors.jl: h5dump
    shows the same groups, attributes, datatypes (string sizes, integer and
    float types) and dataspaces, apart from the storage key."""
    itg, H5 = env
    with H5.h5open(FIXTURE) as f:
        T = H5.read(f, "ITensorv0.1.41", itg.ITensor, device="cpu")
    p = str(tmp_path / "ours.h5")
    with H5.h5open(p, "w") as f:
        H5.write(f, "ITensorv0.1.41", T)
    ref = _header(FIXTURE).replace('GROUP "store"', 'GROUP "storage"')
    assert _header(p) == ref


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_qn_layout_read_by_h5dump(env, tmp_path):
    """QN names as four variable-length strings sorted by name, unused
    slots ("", 0, 0); offsets as N 1-based block coordinates then the
    0-based offset, per block in storage order."""
    itg, H5 = env
    QN, Index = itg.QN, itg.Index
    s = Index([(QN(("Sz", 1), ("Nf", 1)), 1), (QN(("Sz", -1), ("Nf", 1)), 1)], "s")
    t = Index([(QN(("Sz", 1), ("Nf", 1)), 2), (QN(("Sz", -1), ("Nf", 1)), 1)], "t")
    T = itg.random_itensor(QN(), s, t.dag(), rng=np.random.default_rng(3), device="cpu")
    p = str(tmp_path / "qn.h5")
    with H5.h5open(p, "w") as f:
        H5.write(f, "T", T)

    def d(path):
        # the dataset's section of its group's dump ("QN[1]" cannot be
        # named to -d: h5dump reads brackets as a selection)
        group, name = path.rsplit("/", 1)
        out = subprocess.run([H5DUMP, "-g", group, p], check=True, capture_output=True,
                             text=True).stdout
        body = out[out.index(f'DATASET "{name}"'):]
        body = body[body.index("DATA {") + 6:]
        body = re.sub(r"\(\d+\):", " ", body[:body.index("}")])
        return re.findall(r'"[^"]*"|-?\d+', body)

    assert d("/T/inds/index_1/space/QN[1]/names") == ['"Nf"', '"Sz"', '""', '""']
    assert d("/T/inds/index_1/space/QN[1]/vals") == ["1", "1", "0", "0"]
    assert d("/T/inds/index_1/space/QN[1]/mods") == ["1", "1", "0", "0"]
    assert d("/T/inds/index_2/dir") == ["-1"]
    exp = []
    for b, o in T.tensor.blockoffsets.items():
        exp += [str(c) for c in b] + [str(o)]
    assert d("/T/storage/offsets") == exp
    assert d("/T/storage/ndims") == ["2"]


def test_wrong_type_is_refused(env, tmp_path):
    itg, H5 = env
    p = str(tmp_path / "data.h5")
    with H5.h5open(p, "w") as f:
        H5.write(f, "i", itg.Index(2, "i"))
    with H5.h5open(p) as f:
        with pytest.raises(H5.H5Error):
            H5.read(f, "i", itg.ITensor)
