"""The reference's Fourier upsampling rule, restated with numpy's FFT, as the
checker of the device upsampler (mbs_upsample_result).

fft_upsample_phi_arr2d and fft_phi_pad_kernel (src/cuda/GpuDiffraction.cu:
3195-3330, :794-814): forward DFT of the n_in samples of a line; frequency k
placed at k for k <= n_in/2 (an even count's Nyquist term on the positive
side), else at k - n_in; inverse DFT over n_out points scaled by 1/n_in; the
real part.  fft_upsample_theta_arr2d (:3343-3420): each zenith column
extended evenly about both ends (rows t and 2 n_zen - t), upsampled by that
rule to 2 n_zen_out, rows 0..n_zen_out kept."""
import numpy as np


def upsample_line(x, n_out):
    """The rule along the last axis of the real array x."""
    n_in = x.shape[-1]
    X = np.fft.fft(x, axis=-1)
    Y = np.zeros(x.shape[:-1] + (n_out,), complex)
    for k in range(n_in):
        Y[..., (k if k <= n_in // 2 else k - n_in) % n_out] = X[..., k]
    return np.real(np.fft.ifft(Y, axis=-1)) * (n_out / n_in)


def upsample_phi(m, n_out):
    """Along axis 0 (azimuth) of [n_az, ...]."""
    return np.moveaxis(upsample_line(np.moveaxis(m, 0, -1), n_out), -1, 0)


def upsample_theta(m, n_zen_out):
    """Along axis 1 (zenith rows) of [n_az, n_zen+1, ...]."""
    n_zen = m.shape[1] - 1
    ext = np.concatenate([m, m[:, n_zen - 1:0:-1]], axis=1)   # 2 n_zen samples
    up = np.moveaxis(upsample_line(np.moveaxis(ext, 1, -1), 2 * n_zen_out), -1, 1)
    return up[:, :n_zen_out + 1]


def upsample_grid(m, n_az_out, n_zen_out):
    """[n_az, n_zen+1, 16] -> [n_az_out, n_zen_out+1, 16]: zenith, then azimuth."""
    if n_zen_out > m.shape[1] - 1:
        m = upsample_theta(m, n_zen_out)
    if n_az_out > m.shape[0]:
        m = upsample_phi(m, n_az_out)
    return m
