# -*- coding: utf-8 -*-
"""MI355X-native batched GGA forward solve and implicit adjoint (HIP)."""
