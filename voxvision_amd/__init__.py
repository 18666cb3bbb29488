"""voxvision_amd: MI355X (gfx950) ray casting over voxvision octrees."""
