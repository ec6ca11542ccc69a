"""The session's undo levels (src/session.rs:9-51, 139-211) over a stand-in
film on the CPU: generations advance per level and per undo, levels are
trimmed to 64 and to 3 full copies, a stroke whose records the ring has
overwritten is refused (and the stroke levels dropped), and each kind of
level is undone through its own call."""
import pytest

import canvaz_film as czf


class FakeFilm:
    def __init__(self, cap=64):
        self.head, self.cap = 0, cap
        self.gens, self.restored, self.written = [], [], []
        self.state = ["s0"]

    def journal_gen(self, g):
        self.gens.append(g)

    def journal_head(self):
        return self.head, self.cap

    def restore_journal(self, g):
        self.restored.append(g)

    def read_state(self):
        return list(self.state)

    def write_state(self, parts):
        self.written.append(parts)
        return parts != ["bad"]


def test_stroke_levels_restore_their_generation():
    f = FakeFilm()
    u = czf.UndoHistory(f)
    u.begin_stroke()
    f.head = 10
    u.begin_stroke()
    f.head = 30
    assert f.gens == [1, 2]
    u.rollback()
    assert f.restored == [2] and f.gens[-1] == 3
    u.rollback()
    assert f.restored == [2, 1] and f.gens[-1] == 4
    assert not u.can_undo()
    with pytest.raises(czf.FilmError, match="Nothing to undo"):
        u.rollback()


def test_overwritten_stroke_is_refused_and_stroke_levels_dropped():
    f = FakeFilm(cap=64)
    u = czf.UndoHistory(f)
    f.state = ["full-a"]
    u.begin_full()
    u.begin_stroke()          # start head 0
    u.begin_stroke()          # start head 0
    f.head = 65               # 65 records since the last stroke began: more than the ring
    with pytest.raises(czf.FilmError, match="changed too much"):
        u.rollback()
    assert [lv[1] for lv in u.levels] == ["full"]
    assert f.restored == []
    u.rollback()
    assert f.written == [["full-a"]]


def test_head_wraps_modulo_2_32():
    f = FakeFilm(cap=64)
    f.head = 0xFFFFFFF0
    u = czf.UndoHistory(f)
    u.begin_stroke()
    f.head = 0x20  # 48 records later, across the wrap
    u.rollback()
    assert f.restored == [1]


def test_trim_to_64_levels_and_3_full_copies():
    f = FakeFilm()
    u = czf.UndoHistory(f)
    for _ in range(70):
        u.begin_stroke()
    assert len(u.levels) == 64 and u.levels[0][0] == 7
    for k in range(4):
        f.state = [f"full-{k}"]
        u.begin_full()
        u.begin_stroke()
    kinds = [lv[1] for lv in u.levels]
    assert kinds.count("full") == 3
    assert [lv[0] for lv in u.levels] == list(range(72, 79))
    # everything up to and including the oldest surplus full copy is gone
    assert u.levels[0][1] == "journal" and u.levels[1][1] == "full"
    assert u.levels[1][2] == ["full-1"]


def test_failed_full_restore_keeps_the_generation():
    f = FakeFilm()
    u = czf.UndoHistory(f)
    f.state = ["bad"]
    u.begin_full()
    with pytest.raises(czf.FilmError, match="Undo failed"):
        u.rollback()
    assert f.gens == [1]
