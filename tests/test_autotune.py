"""The autotuner's cache without a GPU: set/get, save/load round trip, merge,
and refusal of malformed files (fsim_autotune_set/get/save/load/clear)."""
import pytest

from fsim import autotune
from fsim._lib import FsimError


@pytest.fixture(autouse=True)
def _clean():
    autotune.clear()
    yield
    autotune.clear()


def test_set_get_and_unknown_class():
    assert autotune.get_choice(512, 3, 3, 2, 1) is None
    autotune.set_choice(512, 3, 3, 2, 1, 0, 2)
    assert autotune.get_choice(512, 3, 3, 2, 1) == (0, 2)
    assert autotune.get_choice(512, 3, 3, 2, 0) is None      # another class
    with pytest.raises(FsimError):
        autotune.set_choice(500, 3, 3, 2, 1, 0, 0)             # unsupported size
    with pytest.raises(FsimError):
        autotune.set_choice(512, 3, 3, 2, 1, 2, 0)             # sep not 0/1
    with pytest.raises(FsimError):
        autotune.set_choice(512, 3, 3, 2, 1, 1, 8)             # no such column candidate


def test_save_load_round_trip_and_merge(tmp_path):
    autotune.set_choice(512, 3, 3, 2, 1, 1, 0)
    autotune.set_choice(1024, 3, 5, 3, 1, 0, 3)
    path = tmp_path / "tune.txt"
    autotune.save(path)
    lines = path.read_text().splitlines()
    assert lines[0] == "fsim-autotune 2" and len(lines) == 3
    autotune.clear()
    assert autotune.get_choice(512, 3, 3, 2, 1) is None
    autotune.set_choice(256, 3, 3, 2, 1, 0, 1)                 # kept by the merge
    autotune.load(path)
    assert autotune.get_choice(512, 3, 3, 2, 1) == (1, 0)
    assert autotune.get_choice(1024, 3, 5, 3, 1) == (0, 3)
    assert autotune.get_choice(256, 3, 3, 2, 1) == (0, 1)


@pytest.mark.parametrize("text", ["", "fsim-autotune 3\n", "other 1\n",
                                  "fsim-autotune 1\n512 3 3 2 1 1 9 16 0.5\n", "fsim-autotune 2\n512 3 3 2 1 1 0 16 0.5\n",
                                  "fsim-autotune 2\n512 3 3 2 1 1 0 8 9 16 0.5\n",
                                  "fsim-autotune 1\n512 3 3 2 1 1 0 16\n",
                                  "fsim-autotune 1\n512 3 3 2 7 1 0 16 0.5\n"])
def test_malformed_file_changes_nothing(tmp_path, text):
    autotune.set_choice(128, 3, 3, 2, 1, 0, 1)
    path = tmp_path / "bad.txt"
    path.write_text(text)
    with pytest.raises(FsimError):
        autotune.load(path)
    assert autotune.get_choice(128, 3, 3, 2, 1) == (0, 1)
    assert autotune.get_choice(512, 3, 3, 2, 1) is None


def test_chunk_choice_and_version1_file(tmp_path):
    """Chunking per class: the size default (8 stacks of 512^2 on 2 streams),
    a pinned choice, one launch, the v2 file round trip, and a version 1 file
    (no chunk columns) read with the default chunking."""
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 64) == (8, 2)
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 8) == (8, 1)      # one chunk: one launch
    assert autotune.chunk_plan(1024, 3, 3, 2, 1, 16) == (16, 1)   # no chunking above 512^2 by default
    autotune.set_chunks(512, 3, 3, 2, 1, 16, 3)
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 64) == (16, 3)
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 40) == (16, 3)
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 20) == (16, 2)    # two chunks: two streams
    autotune.set_choice(512, 3, 3, 2, 1, 0, 2)                     # keeps the chunking
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 64) == (16, 3)
    path = tmp_path / "tune.txt"
    autotune.save(path)
    assert path.read_text().splitlines()[1].split()[5:9] == ["0", "2", "16", "3"]
    autotune.set_chunks(512, 3, 3, 2, 1, 0, 1)
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 64) == (64, 1)
    autotune.load(path)
    assert autotune.chunk_plan(512, 3, 3, 2, 1, 64) == (16, 3)
    with pytest.raises(FsimError):
        autotune.set_chunks(512, 3, 3, 2, 1, 8, 5)                 # at most 4 streams
    v1 = tmp_path / "v1.txt"
    v1.write_text("fsim-autotune 1\n256 3 3 2 1 1 0 16 0.5\n")
    autotune.load(v1)
    assert autotune.get_choice(256, 3, 3, 2, 1) == (1, 0)
    assert autotune.chunk_plan(256, 3, 3, 2, 1, 64) == (32, 2)    # 256^2 default: 32 stacks per chunk
