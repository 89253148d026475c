"""Copies the reference's example scene (a .dumb file and the OBJ it names) into
tests/golden/example_scene_0/ so the scene-file tests run where the reference tree
is absent.  The files are data (scene text and mesh), copied byte for byte; the
scene-file tests check the copies against the originals whenever both exist."""
import os
import shutil

SRC = "/root/reference/example_scenes/0"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_scene_0")

if __name__ == "__main__":
    for rel in ("0.dumb", os.path.join("files", "cargo.obj")):
        os.makedirs(os.path.dirname(os.path.join(DST, rel)) or DST, exist_ok=True)
        shutil.copyfile(os.path.join(SRC, rel), os.path.join(DST, rel))
        print("copied", rel)
