"""The reference's interactive controls without a window (lib.rs:41-204 and
WASMInteraction, lib.rs:450-535): the same setters, pointer events and the
redraw that applies them in the reference's order, driven from Python or from
a command script (``Session.execute``; ``python -m lbm_amd.run --script``).

Each ``redraw`` does what one RedrawRequested event does (lib.rs:108-204):
barrier preset, undo, fluid preset, summary and colour-map changes, paint the
stroke in progress, viscosity, then ``rate`` steps (or one in step mode, or
none while paused) and the colour frame, which it returns (None when nothing
was stepped or changed).
"""
from __future__ import annotations

import shlex
from enum import IntEnum

from . import shapes
from .lbm import LBM, ColorMap, SummaryStat, save_ppm

DEFAULT_OMEGA = 1.0 / (0.5 + 0.3)  # lib.rs:38
COMPUTE_PER_RENDER = 15           # lib.rs:17


class FluidPreset(IntEnum):
    """lib.rs:430-437."""
    Equilibrium = 0
    CustomSpeed = 1
    SingleNorth = 2
    SingleOrigin = 3
    SingleEast = 4
    SingleNorthEast = 5


class BarrierPreset(IntEnum):
    """lib.rs:439-445."""
    Welcome = 0
    Tunnel = 1
    Curl = 2
    Chaos = 3


_SINGLE = {FluidPreset.SingleNorth: 1, FluidPreset.SingleOrigin: 4, FluidPreset.SingleEast: 5,
           FluidPreset.SingleNorthEast: 2}


class Session:
    def __init__(self, x: int, y: int, omega: float = DEFAULT_OMEGA, **lbm_kwargs):
        self.lbm = LBM(omega, x, y, **lbm_kwargs)
        self.clicks = shapes.ClickHandler(x, y)
        self.pressed = False
        self.position = (0, 0)
        self.click_type = shapes.ClickType.Draw
        self.click_type_changed = False
        self.paused = False
        self.step_mode = False
        self.step_taken = False
        self.rate = COMPUTE_PER_RENDER
        self.output = SummaryStat.Curl
        self.output_changed = False
        self.color = ColorMap.Jet
        self.color_changed = False
        self.viscosity = None
        self.fluid = FluidPreset.CustomSpeed
        self.fluid_speed = 0.1
        self.fluid_changed = False
        self.barrier = BarrierPreset.Tunnel
        self.barrier_changed = False
        self.undo_count = 0
        self.undo_changed = False
        self.last_frame = None

    def close(self):
        self.lbm.close()

    # -- WASMInteraction setters (lib.rs:452-535) ---------------------------------------
    def set_output(self, stat):
        self.output, self.output_changed = SummaryStat(stat), True

    def set_draw_type(self, click_type: int):
        self.click_type, self.click_type_changed = click_type, True

    def set_color_map(self, cmap):
        self.color, self.color_changed = ColorMap(cmap), True

    def toggle_pause(self):
        self.paused = not self.paused

    def update_compute_rate(self, rate: int):
        self.rate = int(rate)

    def update_viscosity(self, nu: float):
        self.viscosity = float(nu)

    def update_flow_speed(self, speed: float):
        self.fluid, self.fluid_speed, self.fluid_changed = FluidPreset.CustomSpeed, float(speed), True

    def undo(self):
        self.undo_count += 1
        self.undo_changed = True

    def set_step_mode(self):
        self.step_mode = True

    def release_step_mode(self):
        self.step_mode = False

    def take_step(self):
        self.step_taken = True

    def change_fluid_preset(self, preset):
        self.fluid, self.fluid_changed = FluidPreset(preset), True

    def change_barrier_preset(self, preset):
        self.barrier, self.barrier_changed = BarrierPreset(preset), True

    # -- pointer events (lib.rs:80-106), positions in cells ------------------------------
    def cursor_moved(self, x: int, y: int):
        self.position = self.clicks.validate_click((x, y))
        if self.pressed:
            self.clicks.handle_movement(self.position)

    def mouse_input(self, pressed: bool):
        if self.click_type_changed:
            self.clicks.switch_click_type(self.click_type)
            self.click_type_changed = False
        if not pressed:
            self.pressed = False
            self.clicks.handle_release()
        else:
            self.pressed = True
            self.clicks.handle_click(self.position)

    # -- one RedrawRequested (lib.rs:108-204) -------------------------------------------
    def redraw(self):
        g, ch = self.lbm, self.clicks
        barrier_redraw = not ch.current_blob.is_empty() or not ch.current_curve.is_empty()
        if self.barrier_changed:
            g.reset_barrier()
            if self.barrier != BarrierPreset.Tunnel:
                getattr(g, f"{self.barrier.name.lower()}_barrier")()
            ch.clear_barrier()
            barrier_redraw = False
            self.undo_count, self.undo_changed = 0, False
        if self.undo_changed:
            blob = shapes.Blob()
            for _ in range(self.undo_count):
                u = ch.undo()
                if u is not None:
                    blob.join(u)
            if not blob.is_empty():
                g.draw_shape(blob)
            ch.empty_all()
            barrier_redraw = False
            self.undo_count = 0
        if self.fluid_changed:
            if self.fluid == FluidPreset.Equilibrium:
                g.reset_to_equilibrium()
            elif self.fluid == FluidPreset.CustomSpeed:
                g.custom_speed(self.fluid_speed)
            else:
                g.single_cell(_SINGLE[self.fluid])
        if self.output_changed:
            g.set_summary(self.output)
        if self.color_changed:
            g.color_map = self.color
        if barrier_redraw:
            ch.current_blob.join(ch.current_curve)
            g.draw_shape(ch.current_blob)
            ch.update(self.pressed, self.position, self.click_type)
        if self.viscosity is not None:
            g.update_omega_buffer(1.0 / (3.0 * self.viscosity + 0.5))
            self.viscosity = None
        frame = None
        if not self.paused and not self.step_mode:
            g.step(self.rate)
            frame = g.render()
        elif self.step_mode and self.step_taken:
            g.step(1)
            frame = g.render()
            self.step_taken = False
        elif (self.output_changed or barrier_redraw or self.color_changed or self.fluid_changed
              or self.undo_changed or self.barrier_changed):
            frame = g.render()
        self.undo_changed = self.output_changed = self.color_changed = False
        self.fluid_changed = self.barrier_changed = False
        if frame is not None:
            self.last_frame = frame
        return frame

    # -- command scripts -----------------------------------------------------------------
    def execute(self, line: str):
        """One command of a control script; returns the frame a redraw made, if any.

        draw-type line|erase|draw|inactive   press X Y | move X Y | release | click X Y
        undo   preset welcome|tunnel|curl|chaos   fluid <preset>   speed V   viscosity NU
        rate N   pause   step-mode on|off   step   output <stat>   colormap <map>
        frame [N]   save PATH   # comment
        """
        words = shlex.split(line, comments=True)
        if not words:
            return None
        cmd, args = words[0].lower(), words[1:]
        enum_arg = (lambda E: E[args[0].replace("-", " ").title().replace(" ", "")])
        if cmd == "draw-type":
            self.set_draw_type(getattr(shapes.ClickType, args[0].capitalize()))
        elif cmd in ("press", "move", "click"):
            self.cursor_moved(int(args[0]), int(args[1]))
            if cmd != "move":
                self.mouse_input(True)
            if cmd == "click":
                self.mouse_input(False)
        elif cmd == "release":
            self.mouse_input(False)
        elif cmd == "undo":
            self.undo()
        elif cmd == "preset":
            self.change_barrier_preset(enum_arg(BarrierPreset))
        elif cmd == "fluid":
            self.change_fluid_preset(enum_arg(FluidPreset))
        elif cmd == "speed":
            self.update_flow_speed(float(args[0]))
        elif cmd == "viscosity":
            self.update_viscosity(float(args[0]))
        elif cmd == "rate":
            self.update_compute_rate(int(args[0]))
        elif cmd == "pause":
            self.toggle_pause()
        elif cmd == "step-mode":
            self.set_step_mode() if args[0] == "on" else self.release_step_mode()
        elif cmd == "step":
            self.take_step()
        elif cmd == "output":
            self.set_output(enum_arg(SummaryStat))
        elif cmd == "colormap":
            self.set_color_map(enum_arg(ColorMap))
        elif cmd == "frame":
            frame = None
            for _ in range(int(args[0]) if args else 1):
                f = self.redraw()
                frame = f if f is not None else frame
            return frame
        elif cmd == "save":
            if self.last_frame is None:
                raise ValueError("save: no frame rendered yet")
            save_ppm(args[0], self.last_frame)
        else:
            raise ValueError(f"unknown command {cmd!r}")
        return None
