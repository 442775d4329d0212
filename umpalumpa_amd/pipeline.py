"""Image-to-volume chunk loop of the reference's FR harness on one GPU
(examples/fourier_reconstruction/fr.cpp:9-76): for every batch of padded particle images,
FFT (AFFT, R2C) -> crop / centre / normalise / shift (AFP, to (x - 2) x y spatial, i.e. x/2
frequency columns, fr.cpp:166-203) -> insertion into the volume (AFR, fr.cpp:205-228) with
traverse spaces of the batch's orientations times the symmetry operators (fr.cpp:278-312).
All three steps run on the device, on one stream, with no host round trip.  For power-of-two
boxes up to 1024 the FFT and the crop run fused (FourierTransformProcessing: row transforms,
then one kernel for the column transforms and the crop); ``fused=False`` runs the two
operations of the reference instead.
"""
from __future__ import annotations

import numpy as np

from . import fr as F


class ImagePipeline:
    def __init__(self, box: int, batch: int, settings: F.FRSettings, device: int = 0,
                 strategy: int = F.FR_EXACT, symmetries=None, n_table: int = 10000,
                 fused: bool | None = None):
        import torch
        if box % 2:
            raise ValueError("the harness takes even image sizes (fr.cpp:15)")
        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.box, self.batch, self.settings = box, batch, settings
        self.sym = None if symmetries is None else np.asarray(symmetries, np.float64)
        k = 1 if self.sym is None else len(self.sym)
        hx_crop = box // 2                           # (box - 2) / 2 + 1
        fpd = F.fp_desc((box, box), (box - 2, box), batch)
        fps = F.FPSettings(1, 1, 0, 1, 0, 0.0)       # centre, normalise, shift (fr.cpp:192-196)
        if fused is None:
            fused = 4 <= box <= 1024 and box & (box - 1) == 0
        self.fused = fused
        if fused:
            self.fftfp = F.FourierTransformProcessing(device)
            assert self.fftfp.init(fpd, fps), self.fftfp.last_status
        else:
            self.fft = F.FourierTransform(device)
            assert self.fft.init((box, box, 1), batch, F.FFT_FORWARD), self.fft.last_status
            self.fp = F.FourierProcessing(device)
            assert self.fp.init(fpd, fps), self.fp.last_status
        self.desc = F.fr_desc(hx_crop, box, batch, box + 1, batch * k,
                              n_table if settings.interpolation == F.FR_LOOKUP else 0)
        self.op = F.FourierReconstruction(device)
        assert self.op.init(self.desc, settings), self.op.last_status
        self.op.set_strategy(strategy)
        self.table = None
        if settings.interpolation == F.FR_LOOKUP:
            self.table = torch.from_numpy(F.fill_blob_table(n_table, self.desc, settings)).to(
                self.dev)
        self.volume = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=self.dev)
        self.weights = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=self.dev)
        self._spec = None if fused else torch.empty((batch, box, box // 2 + 1),
                                                    dtype=torch.complex64, device=self.dev)
        self._crop = torch.empty((batch, box, hx_crop), dtype=torch.complex64, device=self.dev)

    def spaces(self, rotations):
        """Traverse spaces of the orientations times the symmetry operators, on the device.
        They use the cropped image width box/2, like the harness (fr.cpp:296-305)."""
        sp = F.generate_traverse_spaces(rotations, self.box, self.settings, self.sym)
        return self.torch.from_numpy(sp.view(np.uint8)).to(self.dev)

    def insert(self, images, rotations=None, spaces=None, events=None):
        """images: float32 device tensor [m, box, box] (m <= batch); rotations: [m, 3, 3], or
        `spaces` from spaces() made beforehand.  `events`: four torch events recorded on the
        stream around the FFT, the crop and the insertion (fused: FFT and crop between the
        first two, nothing between the second and third)."""
        m = images.shape[0]
        if m > self.batch:
            raise ValueError("more images than the batch")
        if spaces is None:
            spaces = self.spaces(rotations)
        rec = (lambda i: events[i].record()) if events is not None else (lambda i: None)
        rec(0)
        if self.fused:
            assert self.fftfp.execute(images.contiguous(), m, self._crop), \
                self.fftfp.last_status
            rec(1)
        else:
            assert self.fft.execute(images.contiguous(), self._spec, m), self.fft.last_status
            rec(1)
            assert self.fp.execute(self._spec, m, self._crop, m), self.fp.last_status
        rec(2)
        assert self.op.execute(self._crop[:m], self.volume, self.weights, spaces, self.table), \
            self.op.last_status
        rec(3)

    def synchronize(self):
        self.op.synchronize()
