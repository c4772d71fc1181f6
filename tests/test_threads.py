"""Calls into the C ABI from several host threads (the reference drives its
blocks from OpenMP threads): every entry point holds the library's lock, and
the error message is per thread.  Host-only entry points, so no GPU."""
import ctypes as C
import threading

import numpy as np

from biocfd_amd import hip as H


def test_concurrent_calls_keep_their_own_error_messages():
    lib = H.lib()
    lib.biocfd_last_error.restype = C.c_char_p
    errors = []

    def worker(t):
        info = np.zeros(7, dtype=np.int32)
        try:
            for r in range(300):
                bad = 100000 + 1000 * t + r
                if lib.biocfd_sor_variant_info(bad, H.ptr(info)) == 0:
                    errors.append(f"thread {t}: variant {bad} accepted")
                    return
                msg = lib.biocfd_last_error().decode()
                if str(bad) not in msg:
                    errors.append(f"thread {t}: error of another call: {msg!r} for {bad}")
                    return
                if lib.biocfd_mixed_variant_info(0, H.ptr(info)) != 0 or info[0] <= 0:
                    errors.append(f"thread {t}: valid call failed")
                    return
        except Exception as e:  # pragma: no cover - reported below
            errors.append(f"thread {t}: {e!r}")

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[:5]
