"""Channel mapping and segmentation on the GPU.

Mirrors Mapping.map_channels_segments (python/streamlines/mapping.py:627-776)
and the wrappers it calls -- connect.py, channelheads.py, countlink.py,
label.py, segment.py, hillslopes.py -- with the grids held as device tensors
and every kernel run through libslm_hip.so (slm_amd.hip.map_call).  Host
steps of the reference (flag resets, seed picking, relabelling) are tensor
operations on the same device.

Grids are [nx_padded, ny_padded]: mask bool, uv float32 [..., 2], mapping /
count / link / label int32 tensors holding the reference's uint32 / int32
bits.  The kernel constants are the reference's macro set for the kernel
(pocl.set_compile_options with downup_sign=1, pocl.py:529).
"""
from __future__ import annotations

import numpy as np

from . import hip
from .params import kernel_params

LFA_I32 = -(1 << 31)   # LEFT_FLANK_ADDITION (useful.py:196) as an int32 bit pattern


def pick_seeds(mask=None, map=None, flag=None, pad=0):
    """pick_seeds (useful.py:519-538) on device: pixel (x, y) - pad of the
    selected pixels in argwhere (row-major) order, float32 [N, 2]."""
    import torch
    if mask is None and map is not None:
        sel = (map & int(flag)) != 0
    elif mask is not None and map is None:
        sel = ~mask
    else:
        sel = ~mask & ((map & int(flag)) != 0)
    return (torch.nonzero(sel).to(torch.float32) - float(pad)).contiguous()


class ChannelMapper:
    """State of one mapping pass (the ``data`` arrays of mapping.py on device).

    info      Info (useful.Info) of the cropped grid, flags set
    mask      bool [nxp, nyp] (True = masked)      uv  float32 [nxp, nyp, 2]
    mapping   int32 [nxp, nyp] initial flags (e.g. mid-slope / ridge bits)
    slt, slc  float [nxp, nyp] downstream final fields (map_channels and the
              confluence tie-break); may be None when the channels are given
    """

    def __init__(self, info, mask, uv, mapping, slt=None, slc=None, fp_contract=False, stream=None):
        import torch
        self.info = info
        self.device = mapping.device
        self.mask = mask.to(torch.bool).contiguous()
        self.mask_u8 = self.mask.to(torch.uint8).reshape(-1).contiguous()
        self.uv = uv.to(torch.float32).contiguous()
        self.mapping = mapping.to(torch.int32).contiguous()
        self.slt = None if slt is None else slt.to(torch.float32).contiguous()
        self.slc = None if slc is None else slc.contiguous()
        self.count = None
        self.link = None
        self.label = None
        self.channel_label = None
        self.n_segments = 0
        self.fp_contract = fp_contract
        self.stream = stream
        self.workspace = None
        self.pad = int(info.pad_width)

    # -- helpers ------------------------------------------------------------
    def _params(self, kernel: str, n: int):
        info = self.info
        info.n_seed_points = np.uint32(n)
        p = kernel_params(info, 1, kernel.upper(), self.fp_contract)
        p.n_seed_points = int(n)
        p.seeds_chunk_offset = 0
        return p

    def _run(self, name: str, seeds, **arrays):
        p = self._params(name, seeds.shape[0])
        flat = {k: (v.reshape(-1) if k != "uv" else v.reshape(-1)) for k, v in arrays.items()}
        self.workspace = hip.map_call(name, p, workspace=self.workspace, stream=self.stream, seeds=seeds,
                                      mask=self.mask_u8, **flat)

    def _clear(self, flag):
        m = self.mapping
        m &= ~int(flag)

    # -- mapping.py:695-776 -------------------------------------------------
    def map_channels(self):
        """mapping.py:695-702: mapping = is_channel where slt >= threshold
        and 2*slt >= slc (float64 comparison, as numpy promotes)."""
        slt = self.slt.to(self.device).double()
        slc = self.slc.to(self.device).double()
        sel = (slt >= float(self.info.channel_threshold)) & (slt * 2.0 >= slc)
        self.mapping[sel] = int(self.info.is_channel)

    def connect_channel_pixels(self) -> bool:
        """connect.py:32-67."""
        seeds = pick_seeds(map=self.mapping, flag=self.info.is_channel, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("connect_channels", seeds, uv=self.uv, mapping=self.mapping)
        return True

    def thin_channels(self) -> int:
        """mapping.py:707-719 (stand-in thinning, see slm_thin_channels)."""
        p = self._params("thin_channels", 0)
        need = hip.map_workspace_size(p)
        if self.workspace is None or self.workspace.numel() < need:
            import torch
            self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        return hip.thin_channels(p, self.mapping.reshape(-1), self.workspace, self.stream)

    def map_channel_heads(self) -> bool:
        """channelheads.py:33-79."""
        info = self.info
        thin = int(info.is_thinchannel)
        self.mapping[(self.mapping & thin) == thin] |= int(info.is_channelhead)
        seeds = pick_seeds(mask=self.mask, flag=info.is_channel, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("map_channel_heads", seeds, uv=self.uv, mapping=self.mapping)
        return True

    def count_downchannels(self) -> bool:
        """mapping.py:727-731 + countlink.py:36-80."""
        import torch
        info = self.info
        self.count = torch.zeros_like(self.mapping)
        self.link = torch.zeros_like(self.mapping)
        self._clear(info.is_thinchannel)
        seeds = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_channelhead, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("count_downchannels", seeds, uv=self.uv, mapping=self.mapping, count=self.count, link=self.link)
        return True

    def flag_downchannels(self, do_reset_count=True) -> bool:
        """countlink.py:83-130."""
        info = self.info
        self._clear(info.is_thinchannel)
        if do_reset_count:
            self.count.zero_()
        seeds = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_channelhead, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("flag_downchannels", seeds, mapping=self.mapping, count=self.count, link=self.link)
        return True

    def label_confluences(self) -> bool:
        """mapping.py:737-747 + label.py:33-75: label, then three flag passes."""
        import torch
        info = self.info
        seeds = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_thinchannel, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        slt = self.slt if self.slt is not None else torch.zeros(self.mapping.shape, dtype=torch.float32,
                                                                 device=self.device)
        self._run("label_confluences", seeds, slt=slt, mapping=self.mapping, count=self.count, link=self.link)
        return (self.flag_downchannels() and self.flag_downchannels(do_reset_count=False)
                and self.flag_downchannels(do_reset_count=False))

    def segment_downchannels(self) -> bool:
        """mapping.py:749-757 + segment.py:34-93: segment, relabel 1..n."""
        import torch
        info = self.info
        self.label = torch.zeros_like(self.mapping)
        seeds = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_channelhead, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("segment_downchannels", seeds, mapping=self.mapping, count=self.count, link=self.link,
                  label=self.label)
        pos = self.label > 0
        labels = torch.unique(self.label[pos])
        self.n_segments = int(labels.numel())
        if self.n_segments == 0:
            return False
        self.label[pos] = (torch.searchsorted(labels, self.label[pos]) + 1).to(torch.int32)
        self.channel_label = self.label.clone()
        return True

    def link_hillslopes(self) -> bool:
        """hillslopes.py:32-75."""
        seeds = pick_seeds(mask=self.mask, map=~self.mapping, flag=self.info.is_thinchannel, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("link_hillslopes", seeds, uv=self.uv, link=self.link)
        return True

    def segment_hillslopes(self) -> bool:
        """segment.py:95-146."""
        seeds = pick_seeds(mask=self.mask, map=~self.mapping, flag=self.info.is_thinchannel, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("segment_hillslopes", seeds, uv=self.uv, mapping=self.mapping, label=self.label)
        return True

    def subsegment_flanks(self) -> bool:
        """segment.py:148-243 (four launches) + the sign fix of mapping.py:772-773:
        left-flank labels become -label, right-flank labels +label."""
        info = self.info
        seeds = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_subsegmenthead, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("subsegment_channel_edges", seeds, mapping=self.mapping, link=self.link, label=self.label)
        flank_or_thin = int(info.is_leftflank) | int(info.is_thinchannel)
        seeds = pick_seeds(mask=self.mask, map=~self.mapping, flag=flank_or_thin, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("subsegment_flanks", seeds, uv=self.uv, mapping=self.mapping, label=self.label)
        return self._fix_flanks(seeds)

    def _fix_flanks(self, seeds) -> bool:
        """The last two launches of segment.py:148-243 and the sign fix.
        seeds: those of subsegment_flanks, which fix_right_flanks picks again
        from the same mapping (the flank walk writes labels only)."""
        info = self.info
        self._run("fix_right_flanks", seeds, uv=self.uv, mapping=self.mapping, label=self.label)
        seeds = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_leftflank, pad=self.pad)
        if seeds.shape[0] == 0:
            return False
        self._run("fix_left_flanks", seeds, uv=self.uv, mapping=self.mapping, label=self.label)
        lab = self.label.to(self.device).long()
        neg = lab < 0
        lab[neg] = -(lab[neg] + (1 << 31))
        self.label = lab.to(self.label.dtype)
        return True

    def segment_hillslopes_flanks(self) -> bool:
        """segment_hillslopes() then subsegment_flanks(), with the hillslope
        and flank walks in one launch (slm_segment_hillslopes_flanks): the
        channel edges run first (neither they nor the hillslope walk read
        what the other writes), then one walk per unmasked non-thin pixel,
        then the flank fixes.  The same labels and the same return value as
        the two steps in order."""
        info = self.info
        hs = pick_seeds(mask=self.mask, map=~self.mapping, flag=info.is_thinchannel, pad=self.pad)
        if hs.shape[0] == 0:
            return False
        heads = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_subsegmenthead, pad=self.pad)
        if heads.shape[0] == 0:
            self._run("segment_hillslopes", hs, uv=self.uv, mapping=self.mapping, label=self.label)
            return False
        self._run("subsegment_channel_edges", heads, mapping=self.mapping, link=self.link, label=self.label)
        p = self._params("segment_hillslopes", hs.shape[0])
        self.workspace = hip.map_call("segment_hillslopes_flanks", p, workspace=self.workspace, stream=self.stream,
                                      seeds=hs, mask=self.mask_u8, uv=self.uv.reshape(-1),
                                      mapping=self.mapping.reshape(-1), label=self.label.reshape(-1))
        # the flank seeds: every unmasked pixel lacks one of the two flags
        # (left flanks are never thin), the hillslope seeds among them
        flank_or_thin = int(info.is_leftflank) | int(info.is_thinchannel)
        seeds = pick_seeds(mask=self.mask, map=~self.mapping, flag=flank_or_thin, pad=self.pad)
        return self._fix_flanks(seeds)

    def map_channels_segments(self, do_map_channels_from_scratch=True, do_map_segments=True) -> bool:
        """mapping.py:627-676."""
        if do_map_channels_from_scratch:
            self.map_channels()
        if not self.connect_channel_pixels():
            return False
        self.thin_channels()
        if not self.map_channel_heads():
            return False
        if not self.count_downchannels():
            return False
        if not self.flag_downchannels():
            return False
        if not do_map_segments:
            return True
        return (self.label_confluences() and self.segment_downchannels() and self.link_hillslopes()
                and self.segment_hillslopes_flanks())
