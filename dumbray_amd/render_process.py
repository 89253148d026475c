"""The reference's progressive render loop, BufferedRenderProcess
(BufferedRenderProcess.cu:16-229), over this package's renderers.

A host thread repeats: read the settings under the settings lock; give the back buffer
the target resolution (a change resets the renderer's iterations); unless the renderer
has reached the target iteration count (-1 = no limit) run ``renderer.iterate()`` into
the back buffer, swap front and back when they differ (each buffer then holds one
iteration, as in the reference), and call the iteration / render-completion callbacks;
otherwise call the already-rendered callback.  No window: a negative target resolution
keeps the buffer's size.  Callbacks take the argument they were registered with.
"""
from __future__ import annotations

import threading
import time


class BufferedRenderProcess:
    def __init__(self):
        self._lock = threading.RLock()
        self._cond = threading.Condition(self._lock)
        self._thread = None
        self._kill = False
        self.renderer = None
        self.front = self.back = None
        self.iterations = -1
        self.target_width = self.target_height = -1
        self._callbacks = {"iteration": (None, None), "render": (None, None), "already": (None, None),
                           "error": (None, None)}
        self._render_time = 0.0
        self._start_time = 0.0

    # ------------------------------------------------------------ settings (BufferedRenderProcess.cu:69-100)
    def lockSettings(self):
        self._lock.acquire()

    def unlockSettings(self):
        self._lock.release()

    def setRenderer(self, renderer):
        with self._lock:
            self.renderer = renderer

    def setBuffer(self, buffer):
        with self._lock:
            self.front = self.back = buffer

    def setDoubleBuffers(self, a, b):
        with self._lock:
            self.front, self.back = a, b

    def setTargetIterations(self, n: int):
        with self._lock:
            self.iterations = int(n)

    def setInfinateTargetIterations(self):
        self.setTargetIterations(-1)

    def setTargetResolution(self, width: int, height: int):
        with self._lock:
            self.target_width, self.target_height = int(width), int(height)

    def setIterationCompletionCallback(self, callback, arg=None):
        with self._lock:
            self._callbacks["iteration"] = (callback, arg)

    def setRenderCompletionCallback(self, callback, arg=None):
        with self._lock:
            self._callbacks["render"] = (callback, arg)

    def setAlreadyRenderedCallback(self, callback, arg=None):
        with self._lock:
            self._callbacks["already"] = (callback, arg)

    def setErrorOnIteration(self, callback, arg=None):
        with self._lock:
            self._callbacks["error"] = (callback, arg)

    def synchSettings(self):
        """Returns once the render thread has read the current settings (or at once if it is not running)."""
        with self._cond:
            if self._thread is None:
                return
            self._cond.wait(timeout=60.0)

    def renderTime(self) -> float:
        """Seconds of rendering since the last reset (BufferedRenderProcess.cu:105-107)."""
        with self._lock:
            return self._render_time + ((time.perf_counter() - self._start_time) if self._thread else 0.0)

    # ------------------------------------------------------------ thread
    def start(self):
        with self._lock:
            if self._thread is not None:
                return
            self._kill = False
            self._start_time = time.perf_counter()
            self._thread = threading.Thread(target=self._run, name="BufferedRenderProcess", daemon=True)
            self._thread.start()

    def end(self):
        with self._lock:
            t = self._thread
            if t is None:
                return
            self._kill = True
        t.join()
        with self._lock:
            self._render_time = self.renderTime()
            self._thread = None

    def _call(self, which):
        cb, arg = self._callbacks[which]
        if cb is not None:
            cb(arg)

    def _run(self):
        swap = False
        while True:
            with self._cond:
                if self._kill:
                    break
                renderer = self.renderer
                front, back = (self.back, self.front) if swap else (self.front, self.back)
                target_iterations = self.iterations
                tw, th = self.target_width, self.target_height
                self._cond.notify_all()
            if back is not None and tw > 0 and th > 0 and back.getSize() != (tw, th):
                back.setResolution(tw, th)
                if renderer is not None:
                    renderer.resetIterations()
                    with self._lock:
                        self._render_time, self._start_time = 0.0, time.perf_counter()
            if renderer is None:
                time.sleep(0.008)
                continue
            renderer.setFrameBuffer(back)
            if renderer.iteration() != target_iterations:
                if not renderer.iterate():
                    self._call("error")
                    continue
                if back is not front:
                    swap = not swap
                    renderer.resetIterations()
                    with self._lock:
                        self._render_time, self._start_time = 0.0, time.perf_counter()
                self._call("iteration")
                if renderer.iteration() == target_iterations:
                    self._call("render")
            else:
                self._call("already")
                time.sleep(0.001)
        with self._cond:
            self._cond.notify_all()
