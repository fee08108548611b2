"""The peel (ksm_peel) through the C-ABI from given scores, for GPU tests that
hand the same scores B to the device and to the oracle's peel."""
import torch

from kilosort_amd import _lib
from kilosort_amd import template_matching as tm

DEV = torch.device("cuda")


def gpu_peel(ops, X, U, B, capacity=20000):
    """Runs ksm_peel on B [K, T] (float32, CPU) and the padded batch X [C, T].

    Returns ((st, amps, th_amps, Xres) on the CPU, ctc, nm): ctc [K, K, 2nt+1]
    and nm [K] as the device computed them, for the oracle's peel.
    """
    lib = _lib.load()
    stream = _lib.stream_of(DEV)
    K, P, C = U.shape
    T = X.shape[1]
    nt = ops["nt"]
    Ud = U.to(DEV)
    Wd = ops["wPCA"].to(DEV)
    ctc = tm.prepare_matching(dict(ops, wPCA=Wd), Ud)
    g = tm._grams.get(ctc, Ud, nt, DEV)
    nm = torch.empty((K,), dtype=torch.float32, device=DEV)
    _lib.check(lib.ksm_template_norms(Ud.data_ptr(), K, P, C, nm.data_ptr(), stream), "norms")
    ldbt = (K + 63) // 64 * 64
    Bt = torch.zeros(((T + 127) // 128 * 128, ldbt), dtype=torch.float32, device=DEV)
    perm = g.perm.long()
    Bt[:T, :K] = B.to(DEV)[perm].T
    Xres = X.to(DEV).clone()
    st = torch.empty((capacity, 2), dtype=torch.int64, device=DEV)
    amps = torch.empty((capacity, 1), dtype=torch.float32, device=DEV)
    th = torch.empty((capacity, 1), dtype=torch.float32, device=DEV)
    info = torch.zeros((4,), dtype=torch.int32, device=DEV)
    nb = lib.ksm_peel_workspace(K, T, ops["max_peels"])
    ws = torch.empty((nb,), dtype=torch.uint8, device=DEV)
    _lib.check(lib.ksm_peel(Bt.data_ptr(), ldbt, Xres.data_ptr(), T, Ud.data_ptr(),
                            Wd.data_ptr(), nm.data_ptr(), g.perm.data_ptr(), g.rank.data_ptr(),
                            g.cc.data_ptr(), g.lohi.data_ptr(), g.off.data_ptr(), K, P, C, T,
                            nt, float(ops["Th_learned"]), ops["max_peels"], 0, None, st.data_ptr(),
                            amps.data_ptr(), th.data_ptr(), capacity, info.data_ptr(),
                            ws.data_ptr(), nb, stream), "ksm_peel")
    k, status, iters, _ = (int(v) for v in info.cpu())
    _lib.check(status, "ksm_peel (device)")
    out = (st[:k].cpu(), amps[:k].cpu(), th[:k].cpu(), Xres.cpu())
    ctc_cpu = ctc.cpu().contiguous()
    del ctc, g
    return out, ctc_cpu, nm.cpu()
