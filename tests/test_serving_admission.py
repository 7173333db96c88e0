"""Admission of the continuous-batching server (host logic, no GPU): requests reserve a small
footprint when admitted and their full footprint once their tree gives the voxel counts; a request
that would exceed the budget waits until others release memory, and a request alone always runs."""
import threading
import time
import types

from ngspsr import serving


def _server(budget):
    return types.SimpleNamespace(lock=threading.Condition(), in_use=0, budget=budget)


def _req(fp):
    return types.SimpleNamespace(footprint=fp)


def test_level_footprint_bounds_the_measured_peaks():
    # peaks measured on the serving stream's largest requests (bytes allocated, per-level voxel
    # counts, points): the footprint must not fall below any of them
    measured = [(20.75e9, [5049219, 1777623, 353839, 68457], 898255),
                (14.00e9, [2991759, 1110344, 254233, 54608], 523225),
                (13.09e9, [4245206, 1147260, 229976, 45354], 891526),
                (1.29e9, [224679, 147092, 39910, 8955], 57716),
                (0.71e9, [265084, 54620, 10937, 2191], 457489)]
    for peak, sizes, n in measured:
        fp = serving.level_footprint(sizes, n)
        assert peak <= fp <= 2.5 * peak, (peak, fp)


def test_reserve_waits_for_budget_and_lets_a_lone_request_through():
    srv = _server(budget=10)
    a, b = _req(2), _req(5)
    srv.in_use = a.footprint + b.footprint
    done = threading.Event()

    def grow_a():
        serving.ReconstructionServer._reserve(srv, a, 8)  # 8 + 5 > 10: waits for b
        done.set()

    t = threading.Thread(target=grow_a)
    t.start()
    time.sleep(0.2)
    assert not done.is_set() and a.footprint == 2
    with srv.lock:  # b finishes
        srv.in_use -= b.footprint
        srv.lock.notify_all()
    t.join(5)
    assert done.is_set() and a.footprint == 8 and srv.in_use == 8
    # alone, a request may exceed the budget
    serving.ReconstructionServer._reserve(srv, a, 25)
    assert a.footprint == 25 and srv.in_use == 25
    # a smaller need keeps the reservation
    serving.ReconstructionServer._reserve(srv, a, 3)
    assert a.footprint == 25 and srv.in_use == 25
