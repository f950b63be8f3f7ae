"""Presentation (hpt_present_*): the window loop's hand-off of finished
frames (hala-pathtracer.rs:236-289 presents the final image through the
swapchain after every update()/render()).  Frames arrive in order, each copy
equals the final image hpt_read_image returns at that frame count, and a
loop that presents faster than it acquires gets the newest frame."""
import numpy as np
import pytest

import hpt
import hpt_scenes

pytestmark = pytest.mark.gpu


def _renderer():
    r = hpt.HalaRenderer(96, 64, max_depth=6, rr_depth=3)
    r.set_scene(hpt_scenes.cornell_box())
    r.commit()
    return r


def test_frames_arrive_in_order_and_equal_read_image():
    r = _renderer()
    p = hpt.Presenter(r, slots=3)
    assert p.acquire() is None  # nothing presented yet
    last = 0
    for i in range(6):
        r.render(1 + i % 2)
        p.present()
        r.wait_idle()
        want = r.read_image(hpt.IMAGE_FINAL_RGBA8)
        got = None
        for _ in range(1000):  # the host copy runs on the presenter's stream
            got = p.acquire()
            if got is not None and got[0] == r.frame_index:
                break
        f, img = got
        assert f == r.frame_index and f > last
        assert np.array_equal(img, want)
        last = f
    p.close()
    r.close()


def test_newest_frame_wins_and_full_rings_drop():
    r = _renderer()
    p = hpt.Presenter(r, slots=2)
    finals = {}
    for i in range(8):  # no acquire in between: finished copies are replaced by newer ones
        r.render(1)
        p.present()
        r.wait_idle()
        finals[r.frame_index] = r.read_image(hpt.IMAGE_FINAL_RGBA8)
    # r.wait_idle() waits for the renderer's stream only: the last host copy
    # may still run on the presenter's stream, and acquire does not wait
    f = img = None
    for _ in range(1000):
        got = p.acquire()
        if got is None:  # both slots' copies still in flight
            continue
        assert got[0] in (7, 8)  # with two slots, older copies were overwritten
        f, img = got
        if f == 8:
            break
    assert f == 8 and np.array_equal(img, finals[8])
    # with one slot held by the caller, a present finds at most one slot
    # free: presents beyond it are dropped, never waited for
    for _ in range(3):
        r.render(1)
        p.present()
        # read_image waits for the renderer's stream only, not for the
        # presenter's host copies
        finals[r.frame_index] = r.read_image(hpt.IMAGE_FINAL_RGBA8)
    presented, dropped = p.stats()
    assert presented + dropped == 11 and presented >= 9
    r.wait_idle()
    f2 = None
    for _ in range(1000):
        got = p.acquire()
        if got[0] > 8:
            f2, img2 = got
            break
    assert f2 in (9, 10, 11) and np.array_equal(img2, finals[f2])
    p.close()
    r.close()
