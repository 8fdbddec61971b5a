"""Newton-style State / Control / Contacts containers holding torch tensors.

Newton (the reference's model/state library) is not available; these carry the same
field names the reference reads in `OstrichEngine.step` (core/base_engine.py:303-315,
core/contacts.py:144-176) so call sites look the same.
"""
from dataclasses import dataclass

import torch


@dataclass
class State:
    body_q: torch.Tensor   # [W*B, 7]
    body_qd: torch.Tensor  # [W*B, 6]
    body_f: torch.Tensor   # [W*B, 6]

    @staticmethod
    def from_numpy(d, device):
        return State(*(torch.as_tensor(d[k], dtype=torch.float32, device=device).contiguous()
                       for k in ("body_q", "body_qd", "body_f")))

    def clone(self):
        return State(self.body_q.clone(), self.body_qd.clone(), self.body_f.clone())

    def copy_(self, other):
        self.body_q.copy_(other.body_q)
        self.body_qd.copy_(other.body_qd)
        self.body_f.copy_(other.body_f)
        return self


@dataclass
class Control:
    joint_target_pos: torch.Tensor  # [W*D]
    joint_target_vel: torch.Tensor  # [W*D]

    @staticmethod
    def from_numpy(d, device):
        return Control(*(torch.as_tensor(d[k], dtype=torch.float32, device=device).contiguous()
                         for k in ("joint_target_pos", "joint_target_vel")))


_CONTACT_FIELDS = ("rigid_contact_count", "rigid_contact_point0", "rigid_contact_point1", "rigid_contact_normal",
                   "rigid_contact_shape0", "rigid_contact_shape1", "rigid_contact_margin0", "rigid_contact_margin1")


@dataclass
class Contacts:
    rigid_contact_count: torch.Tensor   # [1] int32
    rigid_contact_point0: torch.Tensor  # [max, 3]
    rigid_contact_point1: torch.Tensor  # [max, 3]
    rigid_contact_normal: torch.Tensor  # [max, 3], from shape0 to shape1
    rigid_contact_shape0: torch.Tensor  # [max] int32, Newton flat shape ids
    rigid_contact_shape1: torch.Tensor  # [max] int32
    rigid_contact_margin0: torch.Tensor # [max]
    rigid_contact_margin1: torch.Tensor # [max]

    @property
    def rigid_contact_max(self):
        return int(self.rigid_contact_shape0.shape[0])

    @staticmethod
    def from_dict(c, device):
        def t(x, dt):
            return torch.as_tensor(x, dtype=dt, device=device).contiguous()
        return Contacts(t(c["count"], torch.int32).reshape(1), t(c["point0"], torch.float32).reshape(-1, 3),
                        t(c["point1"], torch.float32).reshape(-1, 3), t(c["normal"], torch.float32).reshape(-1, 3),
                        t(c["shape0"], torch.int32).reshape(-1), t(c["shape1"], torch.int32).reshape(-1),
                        t(c["margin0"], torch.float32).reshape(-1), t(c["margin1"], torch.float32).reshape(-1))

    def clone(self):
        return Contacts(*(getattr(self, f).clone() for f in _CONTACT_FIELDS))

    def copy_(self, other):
        """Copies another contact set of the same capacity into these tensors (in place,
        so a step captured over them reads the new contacts on replay)."""
        for f in _CONTACT_FIELDS:
            getattr(self, f).copy_(getattr(other, f))
        return self

    def to_dict(self):
        return dict(count=self.rigid_contact_count, max_contacts=self.rigid_contact_max,
                    point0=self.rigid_contact_point0, point1=self.rigid_contact_point1,
                    normal=self.rigid_contact_normal, shape0=self.rigid_contact_shape0,
                    shape1=self.rigid_contact_shape1, margin0=self.rigid_contact_margin0,
                    margin1=self.rigid_contact_margin1)
