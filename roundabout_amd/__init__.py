"""MI355X-native column physics (vertical mixing) for the Roundabout ocean model.

``roundabout_amd.rdb`` drives the HIP library ``lib/librdb_colphys.so``
through its C ABI (``include/rdb_colphys.h``); ``roundabout_amd.synth``
generates the seeded synthetic workloads.
"""
