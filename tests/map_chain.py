"""Test helper: the host steps of Mapping.map_channels_segments
(mapping.py:627-776, segment.py, countlink.py, ...) in numpy around the CPU
oracle (oracle/slm_oracle_map.c), step for step like slm_amd.mapping's
ChannelMapper, so GPU parity tests can compare every stage."""
from __future__ import annotations

import numpy as np

from oracle import medial_axis_ref, oracle
from slm_amd.params import kernel_params
from slm_amd.useful import pick_seeds

STEPS = ("connect", "thin", "heads", "count", "flag", "confluences", "segment", "link_hs",
         "segment_hs", "flanks")


class OracleMapper:
    def __init__(self, info, mask, uv, mapping, slt=None, fp_contract=False):
        self.info = info
        self.mask = np.ascontiguousarray(mask, dtype=bool)
        self.mask_u8 = self.mask.astype(np.uint8).reshape(-1)
        self.uv = np.ascontiguousarray(uv, dtype=np.float32)
        self.mapping = np.ascontiguousarray(mapping, dtype=np.uint32).copy()
        self.slt = None if slt is None else np.ascontiguousarray(slt, dtype=np.float32)
        self.count = self.link = self.label = self.channel_label = None
        self.n_segments = 0
        self.fp_contract = fp_contract
        self.pad = int(info.pad_width)
        self.thin_iters = None

    def params(self, kernel, n):
        self.info.n_seed_points = np.uint32(n)
        p = kernel_params(self.info, 1, kernel.upper(), self.fp_contract)
        p.n_seed_points = int(n)
        return p

    def run(self, name, seeds, **arrays):
        flat = {k: v.reshape(-1) for k, v in arrays.items()}
        return oracle.map_call(name, self.params(name, seeds.shape[0]), seeds=seeds, mask=self.mask_u8, **flat)

    def clear(self, flag):
        self.mapping &= np.uint32(~int(flag) & 0xffffffff)

    def connect_channel_pixels(self):
        s = pick_seeds(map=self.mapping, flag=self.info.is_channel, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("connect_channels", s, uv=self.uv, mapping=self.mapping)
        return True

    def thin_channels(self, medial_axis=True):
        """skeletonize(medial_axis(channel pixels)) (mapping.py:717): the
        medial-axis restatement, then the oracle's thinning of its result."""
        info = self.info
        bits = np.uint32(int(info.is_channel) | int(info.is_interchannel))
        src = self.mapping
        if medial_axis:
            ma = medial_axis_ref.medial_axis((self.mapping & bits) != 0)
            src = np.where(ma, np.uint32(info.is_channel), np.uint32(0)).astype(np.uint32)
        tmp = np.ascontiguousarray(src, dtype=np.uint32).copy()
        self.thin_iters = oracle.map_call("thin_channels", self.params("thin_channels", 0), mapping=tmp.reshape(-1))
        thin = np.uint32(info.is_thinchannel)
        self.mapping[(tmp & thin) == thin] |= thin
        return self.thin_iters

    def map_channel_heads(self):
        info = self.info
        thin = np.uint32(info.is_thinchannel)
        self.mapping[(self.mapping & thin) == thin] |= np.uint32(info.is_channelhead)
        s = pick_seeds(mask=self.mask, flag=info.is_channel, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("map_channel_heads", s, uv=self.uv, mapping=self.mapping)
        return True

    def count_downchannels(self):
        info = self.info
        self.count = np.zeros(self.mapping.shape, np.uint32)
        self.link = np.zeros(self.mapping.shape, np.uint32)
        self.clear(info.is_thinchannel)
        s = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_channelhead, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("count_downchannels", s, uv=self.uv, mapping=self.mapping, count=self.count, link=self.link)
        return True

    def flag_downchannels(self, do_reset_count=True):
        info = self.info
        self.clear(info.is_thinchannel)
        if do_reset_count:
            self.count[:] = 0
        s = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_channelhead, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("flag_downchannels", s, mapping=self.mapping, count=self.count, link=self.link)
        return True

    def label_confluences(self):
        info = self.info
        s = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_thinchannel, pad=self.pad)
        if s.shape[0] == 0:
            return False
        slt = self.slt if self.slt is not None else np.zeros(self.mapping.shape, np.float32)
        self.run("label_confluences", s, slt=slt, mapping=self.mapping, count=self.count, link=self.link)
        return (self.flag_downchannels() and self.flag_downchannels(False) and self.flag_downchannels(False))

    def segment_downchannels(self):
        info = self.info
        self.label = np.zeros(self.mapping.shape, np.uint32)
        s = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_channelhead, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("segment_downchannels", s, mapping=self.mapping, count=self.count, link=self.link,
                 label=self.label)
        lab = self.label.view(np.int32)
        u = np.unique(lab[lab > 0])
        for i, l in enumerate(u):           # segment.py:83-89
            lab[lab == l] = i + 1
        self.n_segments = len(u)
        self.channel_label = self.label.copy()
        return self.n_segments > 0

    def link_hillslopes(self):
        s = pick_seeds(mask=self.mask, map=~self.mapping, flag=self.info.is_thinchannel, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("link_hillslopes", s, uv=self.uv, link=self.link)
        return True

    def segment_hillslopes(self):
        s = pick_seeds(mask=self.mask, map=~self.mapping, flag=self.info.is_thinchannel, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("segment_hillslopes", s, uv=self.uv, mapping=self.mapping, label=self.label)
        return True

    def subsegment_flanks(self):
        info = self.info
        s = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_subsegmenthead, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("subsegment_channel_edges", s, mapping=self.mapping, link=self.link, label=self.label)
        ft = int(info.is_leftflank) | int(info.is_thinchannel)
        s = pick_seeds(mask=self.mask, map=~self.mapping, flag=ft, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("subsegment_flanks", s, uv=self.uv, mapping=self.mapping, label=self.label)
        self.run("fix_right_flanks", s, uv=self.uv, mapping=self.mapping, label=self.label)
        s = pick_seeds(mask=self.mask, map=self.mapping, flag=info.is_leftflank, pad=self.pad)
        if s.shape[0] == 0:
            return False
        self.run("fix_left_flanks", s, uv=self.uv, mapping=self.mapping, label=self.label)
        lab = self.label.view(np.int32).astype(np.int64)
        neg = lab < 0
        lab[neg] = -(lab[neg] + (1 << 31))            # mapping.py:772-773
        self.label = lab.astype(np.int32).view(np.uint32)
        return True

    def step(self, name):
        return {"connect": self.connect_channel_pixels, "thin": self.thin_channels,
                "heads": self.map_channel_heads, "count": self.count_downchannels,
                "flag": self.flag_downchannels, "confluences": self.label_confluences,
                "segment": self.segment_downchannels, "link_hs": self.link_hillslopes,
                "segment_hs": self.segment_hillslopes, "flanks": self.subsegment_flanks}[name]()

    def state(self):
        out = {"mapping": self.mapping.copy()}
        for k in ("count", "link", "label"):
            v = getattr(self, k)
            if v is not None:
                out[k] = v.copy()
        return out


def channels_from_mapping(mapping, info):
    """IS_CHANNEL-only starting grid from a test mapping that marks the axis
    as thin channel (workloads.v_valley)."""
    m = np.zeros(mapping.shape, np.uint32)
    m[(mapping & np.uint32(info.is_thinchannel | info.is_channel)) != 0] = np.uint32(info.is_channel)
    return m
