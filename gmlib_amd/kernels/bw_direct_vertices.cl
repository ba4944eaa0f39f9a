// Direct access, vertex loop: copy the coordinates (VerCrd read, VerDat
// written; the GpuBandwidth pattern of the reference's examples).
VerDat = VerCrd;
