"""The capture guard's deferral list, on the CPU: graphs and events released
while a graph capture is open are kept alive until the outermost capture
closes, whether their owner dies by reference count or in a collection;
outside a capture they are released at once."""

import gc
import weakref

from fvhoe_amd import _lib


class _Graph:
    """Stands in for a captured graph: only its lifetime is observed."""

    def replay(self):
        pass


class _FakeLib:
    def __init__(self):
        self.destroyed = []

    def fvhoe_event_destroy(self, e):
        self.destroyed.append(e.value)


def test_graph_released_inside_a_capture_waits_for_the_capture_to_close():
    g = _Graph()
    alive = weakref.ref(g)
    holder = _lib.GraphHolder(g)
    del g
    with _lib.capture_guard():
        assert _lib.capturing()
        del holder  # reference count reaches zero inside the capture
        assert alive() is not None, "a graph must not be destroyed while a capture is open"
        assert len(_lib._deferred) == 1
    assert not _lib.capturing()
    assert alive() is None and not _lib._deferred


def test_nested_captures_release_at_the_outermost_close():
    g = _Graph()
    alive = weakref.ref(g)
    holder = _lib.GraphHolder(g)
    del g
    with _lib.capture_guard():
        with _lib.capture_guard():
            del holder
        assert alive() is not None  # the outer capture is still open
    assert alive() is None


def test_owner_cycle_is_not_collected_during_a_capture():
    class Owner:
        pass

    def owner():
        g = _Graph()
        o = Owner()
        o.me = o  # a reference cycle, as a solver with its trouble log is
        o.graphs = {"k": (_lib.GraphHolder(g), 3)}
        return o, weakref.ref(g)

    # a cycle already dead is collected on entry, before the capture opens
    o, alive = owner()
    del o
    with _lib.capture_guard():
        assert alive() is None
    # a cycle that dies inside the capture: no automatic collection; an explicit one hands the graph to
    # the deferral list (the collector clears weak references to it first, so identity is checked)
    o, alive = owner()
    gid = id(alive())
    with _lib.capture_guard():
        assert not gc.isenabled()
        del o
        gc.collect()
        assert [id(x) for x in _lib._deferred] == [gid]
    assert not _lib._deferred


def test_outside_a_capture_release_is_immediate():
    g = _Graph()
    alive = weakref.ref(g)
    holder = _lib.GraphHolder(g)
    del g
    del holder
    assert alive() is None and not _lib._deferred


def test_event_destruction_is_deferred_during_a_capture():
    lib = _FakeLib()
    with _lib.capture_guard():
        _lib._destroy_events(lib, 11, 12)
        assert lib.destroyed == []
    assert lib.destroyed == [11, 12]
    _lib._destroy_events(lib, 13)
    assert lib.destroyed == [11, 12, 13]


def test_guard_restores_the_collector():
    was = gc.isenabled()
    with _lib.capture_guard():
        assert not gc.isenabled()
    assert gc.isenabled() == was
