"""Device peak table for the roofline model, keyed by device name
(consts.py:1-12 of the reference holds the same two tables for its NVIDIA
devices).  MI355X: AMD's published dense FP64 rate (vector and matrix) and
HBM3E bandwidth."""

# peak FP64 rate in GFLOP/s
DEV_TO_PEAK_F64_GFLOPS = {
    "AMD Instinct MI355X": 78_600,
}

# peak memory bandwidth in GB/s
DEV_TO_PEAK_BW = {
    "AMD Instinct MI355X": 8_000,
}


# ISA name -> table key, for runtimes that report a generic device name
ARCH_TO_DEVICE = {
    "gfx950": "AMD Instinct MI355X",
}


def lookup_device(name: str, arch: str = "") -> str | None:
    """The table key for a runtime device name (exact, else by substring:
    ROCm reports e.g. ``"AMD Instinct MI355X"`` or a longer marketing name),
    else by the ISA name (``gcnArchName``, e.g. ``"gfx950:sramecc+:xnack-"``)."""
    if name in DEV_TO_PEAK_BW:
        return name
    for key in DEV_TO_PEAK_BW:
        if key.split()[-1] in name:
            return key
    return ARCH_TO_DEVICE.get(arch.split(":")[0])
